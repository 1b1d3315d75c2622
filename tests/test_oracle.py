"""Pins of the CPU oracle (no GPU): golden vectors, reference KATs, closed forms.

Each test is chosen so that a plausible slip in the oracle (a dropped term, a
wrong index or sign, a transposed operand, a wrong step count) fails it.
"""

import os

import numpy as np
import pytest

from oracle import flows
from oracle import native as ON
from oracle import numpy_oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ginterp_vectors.npz")


# ------------------------------------------------------------ interpolation --

def test_interp_matches_reference_golden_vectors():
    """Vectors written by the reference's own GridInterp.cpp (tests/golden)."""
    g = np.load(GOLDEN)
    np.testing.assert_array_equal(O.interp2d_vec(g["v2"], g["p2"]), g["o2"])
    np.testing.assert_array_equal(O.interp3d_vec(g["v3"], g["p3"]), g["o3"])


def test_interp_matches_compiled_reference(rng):
    ref = ON.reference_ginterp()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    v = rng.standard_normal((23, 19))
    p = rng.uniform(-2, 25, (4000, 2))
    out = np.empty(len(p))
    ref.interp2d_vec_inplace(v, p, out)
    np.testing.assert_array_equal(O.interp2d_vec(v, p), out)
    v3 = rng.standard_normal((7, 8, 9))
    p3 = rng.uniform(-1, 10, (4000, 3))
    out3 = np.empty(len(p3))
    ref.interp3d_vec_inplace(v3, p3, out3)
    np.testing.assert_array_equal(O.interp3d_vec(v3, p3), out3)


def test_interp_exact_on_multilinear_fields(rng):
    nx, ny, nz = 9, 7, 5
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    f2 = 0.3 + 1.7 * X - 0.9 * Y + 0.25 * X * Y
    p = rng.uniform(0, [nx - 1 - 1e-9, ny - 1 - 1e-9], (500, 2))
    exp = 0.3 + 1.7 * p[:, 0] - 0.9 * p[:, 1] + 0.25 * p[:, 0] * p[:, 1]
    np.testing.assert_allclose(O.interp2d_vec(f2, p), exp, rtol=1e-12, atol=1e-12)
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    f3 = 1.0 + X - 2 * Y + 3 * Z + 0.5 * X * Y * Z - 0.25 * Y * Z
    p3 = rng.uniform(0, [nx - 1 - 1e-9, ny - 1 - 1e-9, nz - 1 - 1e-9], (500, 3))
    x, y, z = p3.T
    exp3 = 1.0 + x - 2 * y + 3 * z + 0.5 * x * y * z - 0.25 * y * z
    np.testing.assert_allclose(O.interp3d_vec(f3, p3), exp3, rtol=1e-12, atol=1e-12)


def test_interp_boundary_semantics():
    """Zero outside, zero exactly on the upper edge (GridInterp.cpp:126), grid value at the lower edge."""
    v = np.arange(12.0).reshape(4, 3) + 1.0
    pts = np.array([[0, 0], [3, 0], [0, 2], [2.999999, 1.5], [-1e-12, 0], [1, 1], [np.nan, 0]])
    out = O.interp2d_vec(v, pts)
    assert out[0] == v[0, 0]
    assert out[1] == 0.0 and out[2] == 0.0 and out[4] == 0.0 and out[6] == 0.0
    assert out[3] != 0.0
    assert out[5] == v[1, 1]


def test_interp_matches_scipy_regular_grid(rng):
    from scipy.interpolate import RegularGridInterpolator

    v = rng.standard_normal((15, 11))
    rgi = RegularGridInterpolator((np.arange(15.0), np.arange(11.0)), v, method="linear",
                                  bounds_error=False, fill_value=0.0)
    p = rng.uniform(-1, 15, (3000, 2))
    inside_upper = (p[:, 0] == 14) | (p[:, 1] == 10)
    got = O.interp2d_vec(v, p)
    np.testing.assert_allclose(got[~inside_upper], rgi(p)[~inside_upper], rtol=1e-12, atol=1e-12)


def test_c_port_matches_numpy_oracle(rng):
    v = rng.standard_normal((13, 9))
    p = rng.uniform(-1, 14, (2000, 2))
    np.testing.assert_array_equal(ON.cpu_interp2d(v, p), O.interp2d_vec(v, p))
    v3 = rng.standard_normal((6, 7, 5))
    p3 = rng.uniform(-1, 8, (2000, 3))
    np.testing.assert_array_equal(ON.cpu_interp3d(v3, p3), O.interp3d_vec(v3, p3))


def test_grid_field_layouts_follow_reference():
    """2D flat x-fastest (interpolate.py:363); 3D axis swap (interpolate.py:383-385)."""
    nx, ny = 6, 4
    x, y = np.linspace(0, 5, nx), np.linspace(0, 3, ny)
    vel, pts = flows.sample_2d(lambda t, q: np.stack([q[:, 0] + 10 * q[:, 1], -q[:, 1]], 1), 0.0, x, y)
    gf = O.GridField((nx, ny), vel, pts)
    assert gf.v[0].shape == (nx, ny)
    assert gf.v[0][2, 1] == x[2] + 10 * y[1]
    assert gf.deltas == [1.0, 1.0] and gf.mins == [0.0, 0.0]
    n = 5
    a = np.linspace(0, 4, n)
    vel3, pts3 = flows.sample_3d(lambda t, q: np.stack([q[:, 0], 2 * q[:, 1], 3 * q[:, 2]], 1), 0.0, a, a, a)
    g3 = O.GridField((n, n, n), vel3, pts3)
    assert g3.v[0][3, 1, 2] == a[3] and g3.v[1][3, 1, 2] == 2 * a[1] and g3.v[2][3, 1, 2] == 3 * a[2]


# --------------------------------------------------------------- stepping --

def _linear_field(a=0.1, n=41, L=20.0):
    x = np.linspace(0, L, n)
    X, Y = np.meshgrid(x, x)
    pts = np.stack([X.ravel(), Y.ravel()])
    return O.GridField((n, n), a * pts, pts)


def test_euler_kat_reference():
    """tests/test_integrate.py:12-21: v = 0.1 x -> x + h * 0.1 x."""
    p0 = np.arange(1.0, 17.0).reshape(8, 2)
    f = _linear_field()
    x = p0.copy()
    O.euler_step(x, 0.1, f.interpolate)
    np.testing.assert_allclose(x, p0 + 0.1 * p0 * 0.1, rtol=1e-14)


def test_rk4_stability_factor():
    a, h = 0.1, 0.1
    p0 = np.arange(1.0, 17.0).reshape(8, 2)
    f = _linear_field(a)
    x = p0.copy()
    O.rk4_step(x, h, f.interpolate)
    z = a * h
    np.testing.assert_allclose(x, p0 * (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24), rtol=1e-14)


def test_ab2_recurrence():
    a, h = 0.1, 0.1
    f = _linear_field(a)
    x = np.array([[2.0, 3.0]] * 4)
    v = O.ab2_step(x, h, f.interpolate, None)  # Euler first
    np.testing.assert_allclose(x[0], [2.0 * (1 + a * h), 3.0 * (1 + a * h)], rtol=1e-14)
    x1 = x.copy()
    O.ab2_step(x, h, f.interpolate, v)
    np.testing.assert_allclose(x[0], x1[0] + h * (1.5 * a * x1[0] - 0.5 * v[0]), rtol=1e-14)


def test_c_port_advect_matches_oracle():
    vels, pts = [], None
    x, y = flows.grid_axes_2d(48, 24)
    for k in range(6):
        v, pts = flows.sample_2d(flows.double_gyre, 0.1 * k, x, y)
        vels.append(v)
    fields = [O.GridField((48, 24), v, pts) for v in vels]
    x0 = O.make_neighbor_positions(flows.cell_centred_seeds_2d(10, 5), 1e-3)
    for method in ("rk4", "euler", "ab2"):
        _, exp = O.run_window_reference([f.interpolate for f in fields], x0, 0.1, method)
        got = ON.cpu_advect(fields, x0, 0.0, 0.1, 6, method=method, mode=0)
        np.testing.assert_array_equal(got, exp)


# ------------------------------------------------- particles and FTLE --

def test_particle_blocks_reference_values():
    """tests/test_neighboring_particles.py fixtures: [L, R, T, B(, F, Bk)]."""
    p2 = np.array([[0.0, 0.0], [1.0, 0.0], [0.5, 1.0], [0.5, 0.0]])
    np.testing.assert_array_equal(O.delta_right_left(p2, 4), [[1.0, 0.0]])
    np.testing.assert_array_equal(O.delta_top_bottom(p2, 4), [[0.0, 1.0]])
    np.testing.assert_array_equal(O.centroid(p2, 4), [[0.5, 0.25]])
    p3 = np.array([[0, 0, 0], [1, 0, 0], [0.5, 1, 0], [0.5, -1, 0], [0.5, 0, 1], [0.5, 0, -1.0]])
    np.testing.assert_array_equal(O.delta_front_back(p3, 6), [[0.0, 0.0, 2.0]])
    np.testing.assert_array_equal(O.delta_top_bottom(p3, 6), [[0.0, 2.0, 0.0]])


def test_jacobian_of_linear_map(rng):
    for dim in (2, 3):
        M = rng.standard_normal((dim, dim))
        x0 = O.make_neighbor_positions(rng.uniform(-1, 1, (20, dim)), 0.25)
        J = O.flow_map_jacobian(x0 @ M.T, x0)
        np.testing.assert_allclose(J, np.broadcast_to(M, J.shape), rtol=1e-12, atol=1e-12)


def test_ftle_kats(jacobians_2x2, jacobians_3x3):
    """tests/test_ftle.py:17-92."""
    cg = O.cauchy_green_2x2(jacobians_2x2)
    np.testing.assert_allclose(cg, np.einsum("...ij,...kj->...ik", jacobians_2x2, jacobians_2x2), rtol=1e-12)
    np.testing.assert_allclose(O.max_eigenvalue_2x2(cg), [4.0, 4.0], rtol=1e-15)
    np.testing.assert_allclose(O.ftle_2x2(jacobians_2x2, 1.0), [np.log(2)] * 2, rtol=1e-14)
    cg3 = O.cauchy_green_3x3(jacobians_3x3)
    np.testing.assert_allclose(O.max_eigenvalue_3x3(cg3), np.linalg.eigvalsh(cg3)[:, -1], rtol=1e-12)
    np.testing.assert_allclose(O.ftle_3x3(np.repeat(np.eye(3)[None], 2, 0), 1.0), [0.0, 0.0], atol=1e-15)
    # 2D C = F^T F, not F F^T: a non-symmetric F tells them apart in the off-diagonal
    F = np.array([[[1.0, 2.0], [0.0, 1.0]]])
    np.testing.assert_array_equal(O.cauchy_green_2x2(F), [[[1.0, 2.0], [2.0, 5.0]]])


def test_window_arithmetic_saddle_closed_form():
    """Reference window: p snapshots -> p steps, FTLE / ((p-1)|dt|) (ftle_solver.py:99,156,160).

    u = l x, v = -l y is reproduced exactly by bilinear interpolation, so after
    p RK4 steps the x separation grows by R(l h)^p and FTLE = p ln R(l h) / ((p-1) h).
    """
    lam, h, p = 0.5, 0.1, 11
    x = np.linspace(-3, 3, 61)
    vel, pts = flows.sample_2d(lambda t, q: flows.saddle(t, q, lam), 0.0, x, x)
    f = O.GridField((61, 61), vel, pts)
    x0 = O.make_neighbor_positions(np.array([[0.0, 0.5], [0.1, -0.3]]), 1e-3)
    ftle, _ = O.run_window_reference([f.interpolate] * p, x0, h, "rk4")
    z = lam * h
    R = 1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24
    np.testing.assert_allclose(ftle, p * np.log(R) / ((p - 1) * h), rtol=1e-9)


def test_time_linear_blend_exact_for_linear_in_time():
    f0 = lambda q: np.ones_like(q)  # noqa: E731
    f1 = lambda q: 3 * np.ones_like(q)  # noqa: E731
    vat = O.time_linear_velocity([f0, f1], 0.0, 2.0)
    q = np.zeros((3, 2))
    np.testing.assert_allclose(vat(0.5)(q), 1.5)
    np.testing.assert_allclose(vat(2.0)(q), 3.0)
    np.testing.assert_allclose(vat(5.0)(q), 3.0)  # clamped


def test_rgi_field_oracle_closed_forms():
    """RGIField (the reference's GridInterpolator calls, interpolate.py:546-762):
    multilinear fields are exact under linear, tensor cubics under cubic with a direct
    solver, nearest returns grid values, outside -> 0, upper edge inside, NaN -> NaN."""
    from scipy.sparse.linalg import spsolve

    nx, ny, nz = 9, 7, 6
    x, y, z = np.linspace(-1, 2, nx), np.linspace(0, 1, ny), np.linspace(0.5, 1.5, nz)
    X, Y = np.meshgrid(x, y, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel()])
    lin = lambda p, q: 2 + 3 * p - q + 0.5 * p * q  # noqa: E731
    cub = lambda p, q: p**3 - 2 * q**3 + p * q**2 + p**2 * q  # noqa: E731
    vel = np.stack([lin(X, Y).ravel(), cub(X, Y).ravel()])
    q = np.random.default_rng(0).uniform([-1, 0], [2, 1], (200, 2))
    v = O.RGIField((nx, ny), vel, pts, "linear").interpolate(q)
    np.testing.assert_allclose(v[:, 0], lin(q[:, 0], q[:, 1]), atol=1e-13)
    v = O.RGIField((nx, ny), vel, pts, "cubic", solver=spsolve).interpolate(q)
    np.testing.assert_allclose(v[:, 1], cub(q[:, 0], q[:, 1]), atol=1e-12)
    nn = O.RGIField((nx, ny), vel, pts, "nearest").interpolate(np.array([[x[3] + 0.3 * (x[4] - x[3]), y[2]]]))
    assert nn[0, 0] == lin(x[3], y[2])
    edge = O.RGIField((nx, ny), vel, pts, "linear").interpolate(np.array([[2.0, 1.0], [2.0 + 1e-9, 0.5],
                                                                           [np.nan, 0.5]]))
    np.testing.assert_allclose(edge[0, 0], lin(2.0, 1.0), rtol=1e-14)
    assert edge[1, 0] == 0.0 and np.isnan(edge[2, 0])
    # 3D: trilinear exact
    X3, Y3, Z3 = np.meshgrid(x, y, z, indexing="ij")
    f3 = lambda a, b, c: 1 + a - 2 * b + 3 * c + a * b * c  # noqa: E731
    vel3 = np.stack([f3(X3, Y3, Z3).ravel()] * 3)
    pts3 = np.stack([X3.ravel(), Y3.ravel(), Z3.ravel()])
    q3 = np.random.default_rng(1).uniform([-1, 0, 0.5], [2, 1, 1.5], (100, 3))
    v3 = O.RGIField((nx, ny, nz), vel3, pts3, "linear").interpolate(q3)
    np.testing.assert_allclose(v3[:, 2], f3(q3[:, 0], q3[:, 1], q3[:, 2]), atol=1e-13)


# ------------------------------------------------- time-linear window driver --
# run_window_linear / time_linear_velocity have no reference counterpart (the
# reference freezes one snapshot per step, ftle_solver.py:99-102).  They are pinned
# here by closed forms on time-dependent snapshots; each check is shown to reject
# the plausible slips of such a driver (a stage at the wrong time, a wrong divisor,
# an off-by-one snapshot index) by running mutated drivers through it.

def _uniform_field_snapshots(a, b, snap_t0, snap_dt, n_snap):
    """Snapshot k of u(t) = a + b t (spatially uniform) at t_k = snap_t0 + k snap_dt."""
    x, y = np.linspace(-10.0, 10.0, 9), np.linspace(-10.0, 10.0, 7)
    out = []
    for k in range(n_snap):
        def fn(t, q):
            return np.broadcast_to(a + b * t, q.shape).astype(np.float64)

        vel, pts = flows.sample_2d(fn, snap_t0 + k * snap_dt, x, y)
        out.append(O.GridField((9, 7), vel, pts).interpolate)
    return out


def _stretch_field_snapshots(c0, c1, snap_t0, snap_dt, n_snap):
    """Snapshot k of u = c(t) x, v = -c(t) y with c(t) = c0 + c1 t (exact under bilinear)."""
    x = np.linspace(-3.0, 3.0, 13)
    out = []
    for k in range(n_snap):
        c = c0 + c1 * (snap_t0 + k * snap_dt)
        vel, pts = flows.sample_2d(lambda t, q: flows.saddle(t, q, c), 0.0, x, x)
        out.append(O.GridField((13, 13), vel, pts).interpolate)
    return out


def _mutant_window_linear(velocity_at, positions, t0, h, n_steps, kind):
    """run_window_linear with one planted slip (test-only; never used as the oracle)."""
    x = np.array(positions, dtype=np.float64, copy=True)
    x0 = x.copy()
    for i in range(n_steps):
        t = t0 + i * h
        if kind == "k3_at_t_plus_h":
            f1, f2, f3, f4 = velocity_at(t), velocity_at(t + 0.5 * h), velocity_at(t + h), velocity_at(t + h)
        elif kind == "all_stages_at_t":
            f1 = f2 = f3 = f4 = velocity_at(t)
        else:
            f1, f2, f3, f4 = velocity_at(t), velocity_at(t + 0.5 * h), velocity_at(t + 0.5 * h), velocity_at(t + h)
        k1 = f1(x)
        k2 = f2(x + 0.5 * h * k1)
        k3 = f3(x + 0.5 * h * k2)
        k4 = f4(x + h * k3)
        x += (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    T = (n_steps - 1) * abs(h) if kind == "divisor_n_minus_1" else n_steps * abs(h)
    return O.ftle_from_positions(x, x0, T), x


_A = np.array([0.3, -0.2])
_B = np.array([0.1, 0.05])


def test_time_linear_velocity_closed_form_between_and_beyond_snapshots():
    """Snapshots of u(t) = a + b t: the blend is exact at any t inside the series
    (snapshot index floor((t - t0)/dt), weight the remainder) and clamps outside."""
    snaps = _uniform_field_snapshots(_A, _B, 0.5, 0.25, 9)   # t in [0.5, 2.5]
    vat = O.time_linear_velocity(snaps, 0.5, 0.25)
    q = np.array([[0.1, 0.2], [-1.0, 3.0]])
    for t in (0.5, 0.61, 0.75, 1.3, 2.2499, 2.5):
        np.testing.assert_allclose(vat(t)(q), np.broadcast_to(_A + _B * t, q.shape), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(vat(0.1)(q), np.broadcast_to(_A + _B * 0.5, q.shape), rtol=1e-14)
    np.testing.assert_allclose(vat(9.0)(q), np.broadcast_to(_A + _B * 2.5, q.shape), rtol=1e-14)
    # an off-by-one snapshot index or a swapped weight would break the interior values
    shifted = O.time_linear_velocity(snaps, 0.5 - 0.25, 0.25)
    assert not np.allclose(shifted(1.3)(q), _A + _B * 1.3, rtol=1e-6)


def test_window_linear_uniform_field_closed_form_and_mutants():
    """u(t) = a + b t is quadratic-exact for RK4 (Simpson): x(T) = x0 + a T + b (t0 T + T^2/2)
    with stages at t, t+h/2, t+h/2, t+h; k3 at t+h or every stage at t break it."""
    t0, h, n = 0.35, 0.05, 40
    T = n * h
    snaps = _uniform_field_snapshots(_A, _B, 0.0, 0.25, 12)   # t in [0, 2.75] covers [0.35, 2.35]
    vat = O.time_linear_velocity(snaps, 0.0, 0.25)
    x0 = O.make_neighbor_positions(np.array([[0.4, -0.3], [1.1, 0.7], [-2.0, 1.5]]), 1e-3)
    exp = x0 + _A * T + _B * (t0 * T + 0.5 * T * T)
    _, x = O.run_window_linear(vat, x0, t0, h, n, "rk4")
    np.testing.assert_allclose(x, exp, rtol=0, atol=1e-12)
    for kind in ("k3_at_t_plus_h", "all_stages_at_t"):
        _, xm = _mutant_window_linear(vat, x0, t0, h, n, kind)
        assert np.max(np.abs(xm - exp)) > 1e-5, kind
    # Euler: x_n = x0 + h sum_i (a + b t_i); AB2: Euler first, then h(a + b t_i + b h / 2)
    ti = t0 + h * np.arange(n)
    _, xe = O.run_window_linear(vat, x0, t0, h, n, "euler")
    np.testing.assert_allclose(xe, x0 + h * (n * _A + _B * ti.sum()), rtol=0, atol=1e-12)
    _, xa = O.run_window_linear(vat, x0, t0, h, n, "ab2")
    exp_ab2 = x0 + h * (_A + _B * t0) + h * ((n - 1) * _A + _B * (ti[1:].sum() + 0.5 * h * (n - 1)))
    np.testing.assert_allclose(xa, exp_ab2, rtol=0, atol=1e-12)


def test_window_linear_stretch_ftle_closed_form_and_mutants():
    """u = c(t) x, v = -c(t) y, c = c0 + c1 t: the x separation grows by exp(int c dt), so
    FTLE = (1/(n|h|)) int_{t0}^{t0+nh} c dt = c0 + c1 (t0 + n h / 2) (RK4 error ~1e-9 here);
    a divisor of (n-1)|h| or a stage-time slip is off by more than 1e-4 relative."""
    c0, c1, t0, h, n = 0.4, 0.2, 0.3, 0.05, 40
    snaps = _stretch_field_snapshots(c0, c1, 0.0, 0.5, 7)      # t in [0, 3] covers [0.3, 2.3]
    vat = O.time_linear_velocity(snaps, 0.0, 0.5)
    x0 = O.make_neighbor_positions(np.array([[0.05, 0.2], [-0.1, 0.3], [0.2, -0.25]]), 1e-3)
    exact = c0 + c1 * (t0 + 0.5 * n * h)
    ftle, x = O.run_window_linear(vat, x0, t0, h, n, "rk4")
    np.testing.assert_allclose(ftle, exact, rtol=1e-7)
    growth = np.exp(c0 * n * h + c1 * ((t0 + n * h) ** 2 - t0 ** 2) / 2)
    np.testing.assert_allclose(x[:, 0], x0[:, 0] * growth, rtol=1e-7)
    for kind in ("k3_at_t_plus_h", "all_stages_at_t", "divisor_n_minus_1"):
        fm, _ = _mutant_window_linear(vat, x0, t0, h, n, kind)
        assert np.max(np.abs(fm / exact - 1)) > 1e-4, kind
    # the unmutated restatement inside the helper agrees with the oracle
    fr, _ = _mutant_window_linear(vat, x0, t0, h, n, "none")
    np.testing.assert_allclose(fr, ftle, rtol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_c_port_clean_under_sanitizers(dim, tmp_path):
    """The oracle's C restatement built with the address and undefined-behaviour sanitizers
    (``oracle/sanitize_check.c`` driver, serial, -O1, contraction off): interpolation at
    points inside, on the upper edges and outside the grid, and every integrator in frozen
    and time-linear mode with particles that leave the grid, run with no sanitizer report
    and bit-equal to the -O3 OpenMP library on the same inputs."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("gcc absent")
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    exe = str(tmp_path / "sanitize_check")
    build = subprocess.run(
        ["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-fno-fast-math",
         "-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all",
         os.path.join(here, "cpu_path.c"), os.path.join(here, "sanitize_check.c"), "-lm", "-o", exe],
        capture_output=True, text=True)
    if build.returncode != 0 and "asan" in build.stderr.lower():
        pytest.skip("sanitizer runtime unavailable: " + build.stderr[-200:])
    assert build.returncode == 0, build.stderr
    rng = np.random.default_rng(7 + dim)
    shape = (17, 13) if dim == 2 else (9, 8, 7)
    n_snap, n_steps, P, npts = 6, 5, 40, 200
    grid = int(np.prod(shape))
    mins = np.array([0.1, -0.2, 0.3][:dim] + [0.0] * (3 - dim))
    deltas = np.array([0.05, 0.07, 0.06][:dim] + [1.0] * (3 - dim))
    snaps = rng.standard_normal((n_snap, dim, grid))
    ext = np.array(shape) - 1
    # particles: most inside, some near and beyond the edges (they leave the grid)
    x0 = mins[:dim] + deltas[:dim] * ext * rng.uniform(-0.1, 1.1, size=(P, dim))
    # index-space points: inside, exactly on the upper edges, outside on both sides
    pts = rng.uniform(-1.5, ext + 0.5, size=(npts, dim))
    pts[:dim, :] = ext
    pts[dim:2 * dim, :] = ext - 1
    t0, h, snap_t0, snap_dt = 0.0, 0.1, 0.0, 0.1
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array([dim, *shape, *([1] * (3 - dim)), n_snap, P, n_steps, npts], dtype=np.int64).tofile(f)
        np.concatenate([mins, deltas, [snap_t0, snap_dt, t0, h]]).astype(np.float64).tofile(f)
        snaps.astype(np.float64).tofile(f)
        x0.astype(np.float64).tofile(f)
        pts.astype(np.float64).tofile(f)
    out = tmp_path / "out.bin"
    run = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and "runtime error" not in run.stderr and "Sanitizer" not in run.stderr, run.stderr
    got = np.fromfile(out, dtype=np.float64)
    # the same inputs through the -O3 OpenMP library
    lib = ON.cpu_path()
    dp = ON._dp
    v0 = np.ascontiguousarray(snaps[0, 0])
    ref_i = np.empty(npts)
    pc = np.ascontiguousarray(pts)
    if dim == 2:
        lib.cpu_interp2d(dp(v0), shape[0], shape[1], dp(pc), npts, dp(ref_i))
    else:
        lib.cpu_interp3d(dp(v0), *shape, dp(pc), npts, dp(ref_i))
    parts = [ref_i]
    import ctypes

    n = (ctypes.c_int * 3)(*(list(shape) + [1] * (3 - dim)))
    comps = np.ascontiguousarray(snaps)
    for method in range(3):
        for mode in range(2):
            x = np.array(x0, dtype=np.float64, copy=True, order="C")
            assert lib.cpu_advect(dim, n, dp(mins), dp(deltas), dp(comps), n_snap, snap_t0, snap_dt, mode, dp(x), P,
                                  t0, h, n_steps, method, 2) == 0
            parts.append(x.reshape(-1))
    ref = np.concatenate(parts)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)
    assert np.any(ref_i == 0.0) and np.any(ref_i != 0.0)
    # the sanitizers are live: one frozen step past the last snapshot reads beyond the buffer
    hdr = np.fromfile(inp, dtype=np.int64, count=8)
    hdr[6] = n_snap + 1
    raw = bytearray(open(inp, "rb").read())
    raw[:64] = hdr.tobytes()
    bad_in = tmp_path / "bad.bin"
    bad_in.write_bytes(bytes(raw))
    run = subprocess.run([exe, str(bad_in), str(tmp_path / "bad_out.bin")], capture_output=True, text=True,
                         timeout=120)
    assert run.returncode != 0 and "heap-buffer-overflow" in run.stderr
