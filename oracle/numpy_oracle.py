"""NumPy FP64 restatement of pyFTLE's hot path (test infrastructure only).

Every function cites the reference file:line it follows (paths relative to the
pyFTLE source tree ``src/pyftle``).  Arithmetic is written in the same
evaluation order as the reference so that, in FP64, the interpolation matches
the compiled reference ``GridInterp.cpp`` bit for bit and the integrators match
``integrate.py`` to the last ulp.

Pins (checked in ``tests/test_oracle.py``):
* ``interp2d_vec`` / ``interp3d_vec``: bit-exact against the compiled reference
  (``oracle/_ref``) and against committed vectors made from it; exact on
  multilinear fields; SciPy ``RegularGridInterpolator`` cross-check.
* integrators: Euler known answer of ``tests/test_integrate.py:12-21``; RK4
  stability factor ``1+z+z^2/2+z^3/6+z^4/24`` on ``v = a x``; AB2 recurrence.
* particle layout / Jacobian: ``tests/test_neighboring_particles.py`` values and
  the linear-map identity ``J = M``.
* Cauchy-Green / eigenvalue / FTLE: ``tests/conftest.py:189-210`` fixtures with
  ``numpy.linalg.eigvals`` and closed forms (identity -> 0, lambda_max = 4).
* window driver: closed-form saddle ``u = l x, v = -l y`` and the window
  arithmetic of ``ftle_solver.py:99-160``.
* time-linear driver (``time_linear_velocity``, ``run_window_linear``): closed
  forms on time-dependent snapshots -- a uniform ``u(t) = a + b t`` (positions
  exact for RK4, Euler and AB2; the blend exact between snapshots) and a stretch
  ``u = c(t) x`` with ``c`` linear in ``t`` (FTLE = the mean of ``c`` over the
  window, divided by ``n |h|``); mutated drivers (``k3`` at ``t + h``, every
  stage at ``t``, a divisor of ``(n - 1)|h|``) fail these checks.
* ``RGIField``: the reference's ``GridInterpolator`` calls SciPy's
  ``RegularGridInterpolator`` (SciPy 1.15.3 here, the version the reference
  pins); the oracle makes the same calls.  Pinned by closed forms: multilinear
  fields exact for ``linear``, tensor cubics exact for ``cubic`` with a direct
  solver, grid nodes returned exactly by ``nearest``.
"""

from __future__ import annotations

import numpy as np

# ---------------------------------------------------------------------------
# Particle layout: particles.py:153-237 (blocks [L, R, T, B, (F, Bk)])
# ---------------------------------------------------------------------------


def split_blocks(positions: np.ndarray, num_neighbors: int) -> list[np.ndarray]:
    """``np.split(positions, num_neighbors, axis=0)`` as in particles.py:170."""
    return np.split(positions, num_neighbors, axis=0)


def delta_right_left(positions, num_neighbors):
    """particles.py:153-171: ``right - left`` = block 1 - block 0."""
    left, right, *_ = split_blocks(positions, num_neighbors)
    return right - left


def delta_top_bottom(positions, num_neighbors):
    """particles.py:174-192: ``top - bottom`` = block 2 - block 3."""
    _, _, top, bottom, *_ = split_blocks(positions, num_neighbors)
    return top - bottom


def delta_front_back(positions, num_neighbors):
    """particles.py:195-213: ``front - back`` = block 4 - block 5 (3D)."""
    *_, front, back = split_blocks(positions, num_neighbors)
    return front - back


def centroid(positions, num_neighbors):
    """particles.py:216-237: mean over the neighbour blocks."""
    return np.mean(split_blocks(positions, num_neighbors), axis=0)


def make_neighbor_positions(centers: np.ndarray, delta: float) -> np.ndarray:
    """Block-ordered neighbour positions ``(nb*N, dim)`` around ``centers``.

    Order follows file_readers.py:175-180 / particles.py:170,191,212:
    ``[left, right, top, bottom]`` in 2D and ``[left, right, top, bottom,
    front, back]`` in 3D, where left/right differ along x, top/bottom along y
    and front/back along z.
    """
    centers = np.asarray(centers, dtype=np.float64)
    dim = centers.shape[1]
    blocks = []
    for axis in range(dim):
        e = np.zeros(dim)
        e[axis] = delta
        if axis == 0:
            blocks += [centers - e, centers + e]  # left, right
        else:
            blocks += [centers + e, centers - e]  # top, bottom / front, back
    return np.concatenate(blocks, axis=0)


# ---------------------------------------------------------------------------
# Grid interpolation: GridInterp.cpp:15-142 (index space, zero outside)
# ---------------------------------------------------------------------------


def _floor_int(p):
    # GridInterp.cpp:44-46,116-117: static_cast<int>(std::floor(p)).  Values
    # outside int range (and NaN) land out of bounds in both versions.
    with np.errstate(invalid="ignore"):
        f = np.floor(p)
        out = np.where(np.isfinite(f), f, -1.0)
        out = np.clip(out, -(2.0**40), 2.0**40)
    return out.astype(np.int64)


def interp2d_vec(v: np.ndarray, points: np.ndarray, periodic=(False, False)) -> np.ndarray:
    """Bilinear ``interp2d_vec_inplace`` (GridInterp.cpp:93-142).

    ``v`` is ``(nx, ny)`` C-ordered (``idx = x*ny + y``), ``points`` is
    ``(N, 2)`` in index space.  A query is valid iff ``0 <= x0`` and
    ``x0 + 1 < nx`` on every axis (GridInterp.cpp:126); otherwise 0.  The
    ``periodic`` flags are this build's extension: the index wraps modulo the
    axis length instead of the zero fill.
    """
    v = np.ascontiguousarray(v, dtype=np.float64)
    nx, ny = v.shape
    px = points[:, 0].astype(np.float64)
    py = points[:, 1].astype(np.float64)
    x0 = _floor_int(px)
    y0 = _floor_int(py)
    x1 = x0 + 1
    y1 = y0 + 1
    xd = px - x0
    yd = py - y0
    valid = np.ones(px.shape, dtype=bool)
    if periodic[0]:
        x0 = np.mod(x0, nx)
        x1 = np.mod(x1, nx)
    else:
        valid &= (x0 >= 0) & (x1 < nx)
    if periodic[1]:
        y0 = np.mod(y0, ny)
        y1 = np.mod(y1, ny)
    else:
        valid &= (y0 >= 0) & (y1 < ny)
    x0 = np.where(valid, x0, 0)
    x1 = np.where(valid, x1, 0)
    y0 = np.where(valid, y0, 0)
    y1 = np.where(valid, y1, 0)
    v00 = v[x0, y0]
    v10 = v[x1, y0]
    v01 = v[x0, y1]
    v11 = v[x1, y1]
    # GridInterp.cpp:136-137, same evaluation order.
    c = v00 * (1.0 - xd) * (1.0 - yd) + v10 * xd * (1.0 - yd) + v01 * (1.0 - xd) * yd + v11 * xd * yd
    return np.where(valid, c, 0.0)


def interp3d_vec(v: np.ndarray, points: np.ndarray, periodic=(False, False, False)) -> np.ndarray:
    """Trilinear ``interp3d_vec_inplace`` (GridInterp.cpp:15-91).

    ``v`` is ``(X, Y, Z)`` C-ordered, ``points`` ``(N, 3)`` index space;
    x-lerp, then y, then z (GridInterp.cpp:73-84); zero outside (58, 85-87).
    """
    v = np.ascontiguousarray(v, dtype=np.float64)
    X, Y, Z = v.shape
    p = points.astype(np.float64)
    i0 = [_floor_int(p[:, a]) for a in range(3)]
    i1 = [i + 1 for i in i0]
    d = [p[:, a] - i0[a] for a in range(3)]
    dims = (X, Y, Z)
    valid = np.ones(p.shape[0], dtype=bool)
    for a in range(3):
        if periodic[a]:
            i0[a] = np.mod(i0[a], dims[a])
            i1[a] = np.mod(i1[a], dims[a])
        else:
            valid &= (i0[a] >= 0) & (i1[a] < dims[a])
    i0 = [np.where(valid, i, 0) for i in i0]
    i1 = [np.where(valid, i, 0) for i in i1]
    xd, yd, zd = d
    x0, y0, z0 = i0
    x1, y1, z1 = i1
    c00 = v[x0, y0, z0] * (1.0 - xd) + v[x1, y0, z0] * xd
    c01 = v[x0, y0, z1] * (1.0 - xd) + v[x1, y0, z1] * xd
    c10 = v[x0, y1, z0] * (1.0 - xd) + v[x1, y1, z0] * xd
    c11 = v[x0, y1, z1] * (1.0 - xd) + v[x1, y1, z1] * xd
    c0 = c00 * (1.0 - yd) + c10 * yd
    c1 = c01 * (1.0 - yd) + c11 * yd
    c = c0 * (1.0 - zd) + c1 * zd
    return np.where(valid, c, 0.0)


class GridField:
    """One velocity snapshot on a uniform grid, as ``HighPerfInterpolator``.

    Follows interpolate.py:342-424 (layout) and grid_interp.py:30-37,78-79,
    112-121,164-166 (physical -> index ``(p - min) / delta``):

    * 2D: ``velocities.reshape((2, nx, ny), order="F")`` (interpolate.py:363),
      i.e. the flat input is x-fastest (``k = ix + nx*iy``);
    * 3D: ``transpose(velocities.reshape((3, nx, ny, nz)), (0, 2, 1, 3))``
      (interpolate.py:383-385);
    * axes ``linspace(min(coord), max(coord), n)`` (interpolate.py:365-401),
      ``delta = (max - min) / (n - 1)`` (grid_interp.py:34-35).
    """

    def __init__(self, grid_shape, velocities, points, periodic=None):
        grid_shape = tuple(int(n) for n in grid_shape)
        dim = len(grid_shape)
        self.dim = dim
        self.grid_shape = grid_shape
        coordinates = np.asarray(points, dtype=np.float64).reshape((dim, *grid_shape))
        vel = np.asarray(velocities, dtype=np.float64)
        if dim == 2:
            comps = vel.reshape((dim, *grid_shape), order="F")
        elif dim == 3:
            comps = np.transpose(vel.reshape((dim, *grid_shape)), axes=(0, 2, 1, 3))
        else:
            raise ValueError(f"Unsupported grid dimensionality: {dim}")
        self.v = [np.ascontiguousarray(comps[c]) for c in range(dim)]
        self.mins = []
        self.deltas = []
        for a in range(dim):
            axis = np.linspace(np.min(coordinates[a]), np.max(coordinates[a]), grid_shape[a])
            self.mins.append(axis[0])
            self.deltas.append((axis[-1] - axis[0]) / (axis.shape[0] - 1))
        self.periodic = tuple(periodic) if periodic is not None else (False,) * dim

    def to_index(self, pts):
        idx = np.empty_like(pts, dtype=np.float64)
        for a in range(self.dim):
            idx[:, a] = (pts[:, a] - self.mins[a]) / self.deltas[a]
        return idx

    def interpolate(self, pts):
        """interpolate.py:475-543: one component call per axis."""
        pts = np.asarray(pts, dtype=np.float64)
        idx = self.to_index(pts)
        out = np.empty_like(pts, dtype=np.float64)
        f = interp2d_vec if self.dim == 2 else interp3d_vec
        for c in range(self.dim):
            out[:, c] = f(self.v[c], idx, self.periodic)
        return out


# ---------------------------------------------------------------------------
# Integrators: integrate.py:16-109 (step kernels), 161-342 (strategies)
# ---------------------------------------------------------------------------


def euler_step(x, h, f):
    """integrate.py:16-44,187-196: ``x += h * v``."""
    v = f(x)
    x += h * v
    return v


def ab2_step(x, h, f, v_prev):
    """integrate.py:47-78,225-257: Euler when ``v_prev`` is None, else AB2."""
    v = f(x)
    if v_prev is None:
        x += h * v
    else:
        x += h * (1.5 * v - 0.5 * v_prev)
    return v


def rk4_step(x, h, f, f_half=None, f_full=None):
    """integrate.py:291-342 (stages 323-339, update 81-109).

    ``f`` evaluates stage 1; ``f_half`` stages 2-3 and ``f_full`` stage 4 (all
    equal to ``f`` with the reference's frozen field, integrate.py:323-339).
    """
    f_half = f if f_half is None else f_half
    f_full = f if f_full is None else f_full
    k1 = f(x)
    tmp = x.copy()
    tmp += 0.5 * h * k1
    k2 = f_half(tmp)
    tmp = x.copy()
    tmp += 0.5 * h * k2
    k3 = f_half(tmp)
    tmp = x.copy()
    tmp += h * k3
    k4 = f_full(tmp)
    x += (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)


# ---------------------------------------------------------------------------
# Flow-map Jacobian: cauchy_green.py:15-153
# ---------------------------------------------------------------------------


def flow_map_jacobian(positions, initial_positions):
    """``J[:, r, c] = delta_c[:, r] / initial_delta_c[:, c]`` (cauchy_green.py:70-74,140-151)."""
    dim = positions.shape[1]
    nb = 2 * dim
    deltas = [delta_right_left(positions, nb), delta_top_bottom(positions, nb)]
    init = [delta_right_left(initial_positions, nb), delta_top_bottom(initial_positions, nb)]
    if dim == 3:
        deltas.append(delta_front_back(positions, nb))
        init.append(delta_front_back(initial_positions, nb))
    n = positions.shape[0] // nb
    J = np.empty((n, dim, dim))
    for c in range(dim):
        for r in range(dim):
            J[:, r, c] = deltas[c][:, r] / init[c][:, c]
    return J


# ---------------------------------------------------------------------------
# Cauchy-Green, eigenvalue, FTLE: ftle.py:9-233
# ---------------------------------------------------------------------------


def cauchy_green_2x2(F):
    """ftle.py:9-69: ``C = F^T F`` written out as A, B, C, D."""
    C = np.empty_like(F)
    C[:, 0, 0] = F[:, 0, 0] * F[:, 0, 0] + F[:, 1, 0] * F[:, 1, 0]
    C[:, 0, 1] = F[:, 0, 0] * F[:, 0, 1] + F[:, 1, 0] * F[:, 1, 1]
    C[:, 1, 0] = F[:, 0, 1] * F[:, 0, 0] + F[:, 1, 1] * F[:, 1, 0]
    C[:, 1, 1] = F[:, 0, 1] * F[:, 0, 1] + F[:, 1, 1] * F[:, 1, 1]
    return C


def cauchy_green_3x3(F):
    """ftle.py:72-96: ``C = F @ F^T``."""
    return F @ np.transpose(F, (0, 2, 1))


def max_eigenvalue_2x2(C):
    """ftle.py:122-163: ``0.5 * (tr +- sqrt(tr^2 - 4 det))``, max of the two."""
    A, B, Cc, D = C[:, 0, 0], C[:, 0, 1], C[:, 1, 0], C[:, 1, 1]
    trace = A + D
    det = A * D - B * Cc
    with np.errstate(invalid="ignore"):
        disc = np.sqrt(trace**2 - 4 * det)
    return np.maximum(0.5 * (trace + disc), 0.5 * (trace - disc))


def max_eigenvalue_3x3(C):
    """ftle.py:99-119: ``max(np.linalg.eigvals(C))`` (real part for symmetric C)."""
    ev = np.linalg.eigvals(C)
    return np.max(np.real(ev), axis=1)


def ftle_2x2(F, map_period):
    """ftle.py:166-198: ``(1/T) * log(sqrt(lambda_max))``."""
    lam = max_eigenvalue_2x2(cauchy_green_2x2(F))
    with np.errstate(invalid="ignore", divide="ignore"):
        return (1 / map_period) * np.log(np.sqrt(lam))


def ftle_3x3(F, map_period):
    """ftle.py:201-233."""
    lam = max_eigenvalue_3x3(cauchy_green_3x3(F))
    with np.errstate(invalid="ignore", divide="ignore"):
        return (1 / map_period) * np.log(np.sqrt(lam))


def ftle_from_positions(final_positions, initial_positions, map_period):
    """ftle_solver.py:122-161: Jacobian then 2D/3D FTLE by neighbour count."""
    J = flow_map_jacobian(final_positions, initial_positions)
    if final_positions.shape[1] == 2:
        return ftle_2x2(J, map_period)
    return ftle_3x3(J, map_period)


# ---------------------------------------------------------------------------
# Window drivers
# ---------------------------------------------------------------------------


def run_window_reference(snapshot_fields, positions, dt, integrator="rk4"):
    """Reference window semantics (ftle_solver.py:93-161, integrate.py:323-342).

    ``snapshot_fields`` is a list of ``p`` callables (one per snapshot file).
    Step ``i`` uses snapshot ``i`` frozen for all stages; ``p`` steps are taken
    and the FTLE divides by ``(p - 1) * |dt|`` (ftle_solver.py:156,160).
    Returns ``(ftle, final_positions)``.
    """
    x = np.array(positions, dtype=np.float64, copy=True)
    x0 = x.copy()
    v_prev = None
    p = len(snapshot_fields)
    for i in range(p):
        f = snapshot_fields[i]
        if integrator == "rk4":
            rk4_step(x, dt, f)
        elif integrator == "euler":
            euler_step(x, dt, f)
        elif integrator == "ab2":
            v_prev = ab2_step(x, dt, f, v_prev).copy()
        else:
            raise ValueError(integrator)
    T = (p - 1) * abs(dt)
    return ftle_from_positions(x, x0, T), x


def time_linear_velocity(fields, snap_t0, snap_dt):
    """Velocity at time ``t`` by linear interpolation between snapshots.

    This build's extension (no reference counterpart): snapshot ``k`` sits at
    ``snap_t0 + k*snap_dt``; ``k = floor((t - t0)/dt)`` is clamped to
    ``[0, n-2]`` and the weight to ``[0, 1]``.
    """
    n = len(fields)

    def at(t):
        s = (t - snap_t0) / snap_dt
        k = int(np.floor(s))
        k = min(max(k, 0), n - 2) if n > 1 else 0
        w = min(max(s - k, 0.0), 1.0) if n > 1 else 0.0

        def f(x):
            if n == 1 or w == 0.0:
                return fields[k](x)
            return (1.0 - w) * fields[k](x) + w * fields[k + 1](x)

        return f

    return at


def run_window_linear(velocity_at, positions, t0, h, n_steps, integrator="rk4"):
    """Time-continuous window: RK4 stages at ``t, t+h/2, t+h/2, t+h``.

    ``velocity_at(t)`` returns a callable ``f(x)``.  The FTLE divides by
    ``n_steps * |h|`` (the exact integration time).  This build's
    parameterisation for time-linear snapshots and closed-form flows.
    """
    x = np.array(positions, dtype=np.float64, copy=True)
    x0 = x.copy()
    v_prev = None
    for i in range(n_steps):
        t = t0 + i * h
        if integrator == "rk4":
            rk4_step(x, h, velocity_at(t), velocity_at(t + 0.5 * h), velocity_at(t + h))
        elif integrator == "euler":
            euler_step(x, h, velocity_at(t))
        elif integrator == "ab2":
            v_prev = ab2_step(x, h, velocity_at(t), v_prev).copy()
        else:
            raise ValueError(integrator)
    T = n_steps * abs(h)
    return ftle_from_positions(x, x0, T), x


# ---------------------------------------------------------------------------
# GridInterpolator: interpolate.py:546-762 (SciPy RegularGridInterpolator)
# ---------------------------------------------------------------------------
class RGIField:
    """One snapshot as the reference's ``GridInterpolator`` sees it after an update
    with points: each component reshaped to ``grid_shape`` in C order
    (interpolate.py:626-660), axes ``linspace(points[a].min(), points[a].max(), n)``
    (interpolate.py:620-625), ``RegularGridInterpolator(..., method, bounds_error=False,
    fill_value=0.0)`` per component; ``interpolate`` fills an ``(M, dim)`` array
    (interpolate.py:680-700).  ``solver``: ``None`` = SciPy's default for spline
    methods (an iterative solver), or e.g. ``scipy.sparse.linalg.spsolve`` for the
    exact spline.
    """

    def __init__(self, grid_shape, velocities, points, method="linear", solver=None):
        from scipy.interpolate import RegularGridInterpolator

        grid_shape = tuple(int(n) for n in grid_shape)
        pts = np.asarray(points, dtype=np.float64)
        vel = np.asarray(velocities, dtype=np.float64)
        self.dim = len(grid_shape)
        self.grid = tuple(np.linspace(pts[a].min(), pts[a].max(), grid_shape[a]) for a in range(self.dim))
        kw = {}
        if solver is not None:
            kw["solver"] = solver
        self.comps = [RegularGridInterpolator(self.grid, vel[c].reshape(grid_shape), method=method,
                                              bounds_error=False, fill_value=0.0, **kw)
                      for c in range(self.dim)]

    def interpolate(self, new_points):
        new_points = np.asarray(new_points, dtype=np.float64)
        out = np.empty_like(new_points)
        for c in range(self.dim):
            out[:, c] = self.comps[c](new_points)
        return out


# ---------------------------------------------------------------------------
# Scattered-data strategies: interpolate.py:188-294 (SciPy LinearNDInterpolator /
# NearestNDInterpolator on the data points)
# ---------------------------------------------------------------------------
class ScatteredField:
    """The reference's ``CubicInterpolator`` (2D) / ``LinearInterpolator`` /
    ``NearestNeighborInterpolator`` after an
    update: ``u + i v`` through one complex interpolator (interpolate.py:216-220, 272-273),
    ``w`` through a second real one in 3D (interpolate.py:226-229, 274-277); ``interpolate``
    fills ``(M, dim)`` (interpolate.py:231-246, 279-294).  The reference fills ``w`` only
    when exactly three points are queried (it tests ``len(new_points) == 3``); the oracle
    always fills it."""

    def __init__(self, points, velocities, method="linear"):
        from scipy.interpolate import CloughTocher2DInterpolator, LinearNDInterpolator, NearestNDInterpolator

        pts = np.asarray(points, dtype=np.float64)
        vel = np.asarray(velocities, dtype=np.float64)
        self.dim = pts.shape[0]
        cls = {"linear": LinearNDInterpolator, "nearest": NearestNDInterpolator,
               "cubic": CloughTocher2DInterpolator}[method]  # cubic: interpolate.py:136-155, 2D
        self.uv = cls(pts.T, vel[0] + 1j * vel[1])
        self.w = None
        if self.dim == 3:
            self.w = (LinearNDInterpolator(self.uv.tri, vel[2]) if method == "linear"
                      else NearestNDInterpolator(pts.T, vel[2]))

    def interpolate(self, new_points):
        new_points = np.asarray(new_points, dtype=np.float64)
        out = np.empty_like(new_points)
        uv = self.uv(new_points)
        out[:, 0] = uv.real
        out[:, 1] = uv.imag
        if self.w is not None:
            out[:, 2] = self.w(new_points)
        return out
