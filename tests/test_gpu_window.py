"""GPU parity of the fused window kernel and the drop-in API against the oracle.

FP64 tolerance: positions bit-exact with the oracle (same operation order,
no contraction); FTLE to 1e-12 relative (device log/sqrt, 3D eigen-solver).
FP32 tolerance: FTLE max abs error <= 1e-3 * (max - min) of the FP64 field.
"""

import numpy as np
import pytest

from oracle import numpy_oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyftle_amd import _native

    _native.load()


def _field_2d(vels, pts, grid_shape, precision, pair=False, dt_snap=0.1):
    from pyftle_amd import _native as N
    from pyftle_amd.field import DeviceField, GridGeometry

    geom = GridGeometry.from_reference(grid_shape, pts)
    f = DeviceField(geom, dtype=precision, layout=N.LAYOUT_PAIR if pair else N.LAYOUT_SINGLE,
                    ring_slots=len(vels), n_snapshots=len(vels), t0=0.0, dt_snap=dt_snap)
    for k, v in enumerate(vels):
        f.pack(k, v)
    return f


def _rel_range_err(got, exp):
    rng = np.nanmax(exp) - np.nanmin(exp)
    return np.nanmax(np.abs(got - exp)) / rng


@pytest.mark.parametrize("method", ["rk4", "euler", "ab2"])
def test_window_reference_mode_fp64_bit_exact(method):
    from pyftle_amd import window as W

    vels, pts = H.double_gyre_window(nx=64, ny=32, n_snap=11)
    x0 = H.seeds_2d(16, 8)
    fields = H.grid_fields((64, 32), vels, pts)
    exp_ftle, exp_pos = O.run_window_reference([f.interpolate for f in fields], x0, 0.1, method)
    f = _field_2d(vels, pts, (64, 32), "float64")
    pos = torch.as_tensor(x0, device="cuda")
    out = torch.empty_like(pos)
    ftle = W.advect_ftle(f, pos, method=method, time_mode="frozen", n_steps=11, h=0.1,
                         map_period=10 * 0.1, positions_out=out)
    np.testing.assert_array_equal(out.cpu().numpy(), exp_pos)
    np.testing.assert_allclose(ftle.cpu().numpy(), exp_ftle, rtol=1e-12, atol=1e-14)


def test_window_reference_mode_fp32_tolerance():
    from pyftle_amd import window as W

    vels, pts = H.double_gyre_window(nx=128, ny=64, n_snap=21)
    x0 = H.seeds_2d(40, 20)
    fields = H.grid_fields((128, 64), vels, pts)
    exp_ftle, _ = O.run_window_reference([f.interpolate for f in fields], x0, 0.1, "rk4")
    f = _field_2d(vels, pts, (128, 64), "float32")
    pos = torch.as_tensor(x0, device="cuda")
    ftle = W.advect_ftle(f, pos, method="rk4", time_mode="frozen", n_steps=21, h=0.1, map_period=2.0)
    err = _rel_range_err(ftle.cpu().numpy(), exp_ftle)
    print(f"[fp32 frozen 2D] max|err|/range = {err:.3e}")
    assert err <= 1e-3


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_window_time_linear_2d(precision):
    from pyftle_amd import window as W

    vels, pts = H.double_gyre_window(nx=96, ny=48, n_snap=21)
    x0 = H.seeds_2d(24, 12)
    fields = H.grid_fields((96, 48), vels, pts)
    vat = O.time_linear_velocity([f.interpolate for f in fields], 0.0, 0.1)
    exp_ftle, exp_pos = O.run_window_linear(vat, x0, 0.0, 0.05, 40, "rk4")
    f = _field_2d(vels, pts, (96, 48), precision, pair=True)
    pos = torch.as_tensor(x0, device="cuda")
    out = torch.empty_like(pos)
    ftle = W.advect_ftle(f, pos, method="rk4", time_mode="linear", n_steps=40, h=0.05,
                         t_start=0.0, map_period=2.0, positions_out=out)
    if precision == "float64":
        np.testing.assert_array_equal(out.cpu().numpy(), exp_pos)
        np.testing.assert_allclose(ftle.cpu().numpy(), exp_ftle, rtol=1e-12, atol=1e-14)
    else:
        err = _rel_range_err(ftle.cpu().numpy(), exp_ftle)
        print(f"[fp32 linear 2D] max|err|/range = {err:.3e}")
        assert err <= 1e-3


@pytest.mark.parametrize("method", ["rk4", "euler", "ab2"])
def test_window_time_linear_closed_forms(method):
    """The time-linear kernel against closed forms, not the oracle: snapshots of the
    uniform u(t) = a + b t (positions exact for RK4 as Simpson's rule, and for the
    Euler/AB2 sums) and of the stretch u = c(t) x, v = -c(t) y (RK4 FTLE = mean of c)."""
    from oracle import flows
    from pyftle_amd import window as W

    a, b = np.array([0.3, -0.2]), np.array([0.1, 0.05])
    t0, h, n = 0.35, 0.05, 40
    T = n * h
    x, y = np.linspace(-10.0, 10.0, 9), np.linspace(-10.0, 10.0, 7)
    vels = []
    for k in range(12):
        v, pts = flows.sample_2d(lambda t, q: np.broadcast_to(a + b * t, q.shape).astype(np.float64),
                                 0.25 * k, x, y)
        vels.append(v)
    f = _field_2d(vels, pts, (9, 7), "float64", pair=True, dt_snap=0.25)
    x0 = O.make_neighbor_positions(np.array([[0.4, -0.3], [1.1, 0.7], [-2.0, 1.5]]), 1e-3)
    pos = torch.as_tensor(x0, device="cuda")
    out = torch.empty_like(pos)
    W.advect_ftle(f, pos, method=method, time_mode="linear", n_steps=n, h=h, t_start=t0,
                  map_period=T, positions_out=out)
    ti = t0 + h * np.arange(n)
    if method == "rk4":
        exp = x0 + a * T + b * (t0 * T + 0.5 * T * T)
    elif method == "euler":
        exp = x0 + h * (n * a + b * ti.sum())
    else:
        exp = x0 + h * (a + b * t0) + h * ((n - 1) * a + b * (ti[1:].sum() + 0.5 * h * (n - 1)))
    np.testing.assert_allclose(out.cpu().numpy(), exp, rtol=0, atol=1e-12)
    if method != "rk4":
        return
    c0, c1 = 0.4, 0.2
    g = np.linspace(-3.0, 3.0, 13)
    vels = [flows.sample_2d(lambda t, q: flows.saddle(t, q, c0 + c1 * 0.5 * k), 0.0, g, g)[0] for k in range(7)]
    pts = flows.sample_2d(flows.saddle, 0.0, g, g)[1]
    f = _field_2d(vels, pts, (13, 13), "float64", pair=True, dt_snap=0.5)
    x0 = O.make_neighbor_positions(np.array([[0.05, 0.2], [-0.1, 0.3], [0.2, -0.25]]), 1e-3)
    ftle = W.advect_ftle(f, torch.as_tensor(x0, device="cuda"), method="rk4", time_mode="linear",
                         n_steps=n, h=h, t_start=0.3, map_period=n * h)
    np.testing.assert_allclose(ftle.cpu().numpy(), c0 + c1 * (0.3 + 0.5 * n * h), rtol=1e-7)


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_window_3d_periodic_abc(precision):
    from pyftle_amd import _native as N
    from pyftle_amd import window as W
    from pyftle_amd.field import DeviceField, GridGeometry

    n = 24
    vels, pts = H.abc_window(n=n)
    per = (True, True, True)
    fields = H.grid_fields((n, n, n), vels, pts, periodic=per)
    x0 = H.seeds_3d(5)
    steps, h = 60, 0.05
    exp_ftle, exp_pos = O.run_window_linear(lambda t: fields[0].interpolate, x0, 0.0, h, steps, "rk4")
    geom = GridGeometry.from_reference((n, n, n), pts, periodic=per)
    f = DeviceField(geom, dtype=precision, layout=N.LAYOUT_SINGLE, ring_slots=1, n_snapshots=1)
    f.pack(0, vels[0])
    pos = torch.as_tensor(x0, device="cuda")
    out = torch.empty_like(pos)
    ftle = W.advect_ftle(f, pos, method="rk4", time_mode="steady", n_steps=steps, h=h,
                         map_period=steps * h, positions_out=out)
    if precision == "float64":
        np.testing.assert_array_equal(out.cpu().numpy(), exp_pos)
        np.testing.assert_allclose(ftle.cpu().numpy(), exp_ftle, rtol=1e-10, atol=1e-13)
    else:
        err = _rel_range_err(ftle.cpu().numpy(), exp_ftle)
        print(f"[fp32 steady 3D periodic] max|err|/range = {err:.3e}")
        assert err <= 1e-3


def test_ftle_solver_array_source_matches_oracle():
    from pyftle_amd import NeighboringParticles
    from pyftle_amd.data_source import ArrayBatchSource
    from pyftle_amd.ftle_solver import FTLESolver
    from pyftle_amd.integrate import create_integrator
    from pyftle_amd.interpolate import create_interpolator

    vels, pts = H.double_gyre_window(nx=64, ny=32, n_snap=11)
    x0 = H.seeds_2d(16, 8)
    fields = H.grid_fields((64, 32), vels, pts)
    exp_ftle, exp_pos = O.run_window_reference([f.interpolate for f in fields], x0, 0.1, "rk4")
    interp = create_interpolator("grid", (64, 32), precision="float64")
    integ = create_integrator("RK4", interp)
    parts = NeighboringParticles(x0.copy())
    src = ArrayBatchSource(vels, pts, parts, 0.1)
    ftle = FTLESolver(src, integ).run()
    np.testing.assert_allclose(ftle, exp_ftle, rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(parts.positions, exp_pos)


@pytest.mark.parametrize("slots", [1, 2, 5])
@pytest.mark.parametrize("on_device", [False, True])
def test_ftle_solver_chunked_ring_matches_single_launch(slots, on_device):
    """A ring smaller than the window (several launches, packing on a second stream
    overlapped with the previous launch) gives the oracle's field; device snapshots
    written on the compute stream just before the window are seen whole."""
    from pyftle_amd import NeighboringParticles
    from pyftle_amd.data_source import ArrayBatchSource
    from pyftle_amd.ftle_solver import FTLESolver
    from pyftle_amd.integrate import create_integrator
    from pyftle_amd.interpolate import HighPerfInterpolator

    vels, pts = H.double_gyre_window(nx=64, ny=32, n_snap=13)
    x0 = H.seeds_2d(12, 6)
    for method in ("rk4", "ab2", "euler"):
        fields = H.grid_fields((64, 32), vels, pts)
        exp_ftle, exp_pos = O.run_window_reference([f.interpolate for f in fields], x0, 0.1, method)
        small = HighPerfInterpolator((64, 32), precision="float64",
                                     max_ring_bytes=slots * 64 * 32 * 2 * 8)
        src_vels = vels
        if on_device:
            # filled by a kernel on the current stream, not yet complete when run() starts
            src_vels = [torch.zeros(v.shape, dtype=torch.float64, device="cuda") for v in vels]
            for d, v in zip(src_vels, vels):
                d.add_(torch.as_tensor(v).to("cuda", non_blocking=True))
        parts = NeighboringParticles(x0.copy())
        ftle = FTLESolver(ArrayBatchSource(src_vels, pts, parts, 0.1), create_integrator(method, small)).run()
        np.testing.assert_allclose(ftle, exp_ftle, rtol=1e-12, atol=1e-14)
        np.testing.assert_array_equal(parts.positions, exp_pos)


@pytest.mark.parametrize("name", ["euler", "rk4", "ab2"])
def test_per_step_integrators_match_oracle(name):
    """Integrator.integrate(h, particles): one kernel step per call (integrate.py:161-342)."""
    from pyftle_amd import NeighboringParticles
    from pyftle_amd.integrate import create_integrator
    from pyftle_amd.interpolate import create_interpolator

    vels, pts = H.double_gyre_window(nx=64, ny=32, n_snap=4)
    x0 = H.seeds_2d(8, 4)
    interp = create_interpolator("grid", (64, 32), precision="float64")
    integ = create_integrator(name, interp)
    parts = NeighboringParticles(x0.copy())
    xo = x0.copy()
    v_prev = None
    for i, v in enumerate(vels):
        interp.update(v, pts if i == 0 else None)
        integ.integrate(0.1, parts)
        f = O.GridField((64, 32), v, pts).interpolate
        if name == "rk4":
            O.rk4_step(xo, 0.1, f)
        elif name == "euler":
            O.euler_step(xo, 0.1, f)
        else:
            v_prev = O.ab2_step(xo, 0.1, f, v_prev).copy()
        np.testing.assert_array_equal(parts.positions, xo)


def test_euler_kat_linear_field():
    """tests/test_integrate.py:12-21: v = 0.1 x gives x + h*0.1*x (bilinear is exact)."""
    from pyftle_amd import NeighboringParticles
    from pyftle_amd.integrate import create_integrator
    from pyftle_amd.interpolate import create_interpolator

    x = np.linspace(0, 20, 21)
    X, Y = np.meshgrid(x, x)
    pts = np.stack([X.ravel(), Y.ravel()])
    vel = 0.1 * pts
    interp = create_interpolator("grid", (21, 21), precision="float64")
    interp.update(vel, pts)
    p0 = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0], [7.0, 8.0],
                   [9.0, 10.0], [11.0, 12.0], [13.0, 14.0], [15.0, 16.0]])
    parts = NeighboringParticles(p0.copy())
    create_integrator("euler", interp).integrate(0.1, parts)
    np.testing.assert_allclose(parts.positions, p0 + 0.1 * p0 * 0.1, rtol=1e-14)
    parts = NeighboringParticles(p0.copy())
    create_integrator("rk4", interp).integrate(0.1, parts)
    z = 0.01
    np.testing.assert_allclose(parts.positions, p0 * (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24), rtol=1e-14)


@pytest.mark.parametrize("method", ["rk4", "euler"])
def test_file_experiment_batched_equals_window_by_window(tmp_path, method):
    """All windows of a file experiment in one persistent launch give the FTLE fields of
    FTLESolver run window by window, bit for bit (forward and backward time)."""
    from scipy.io import savemat

    from pyftle_amd.experiment import FileFTLEManager, load_config

    nx, ny = 64, 32
    vels, pts = H.double_gyre_window(nx=nx, ny=ny, n_snap=12)
    cx, cy = pts[0].reshape(ny, nx), pts[1].reshape(ny, nx)
    savemat(tmp_path / "c.mat", {"coordinate_x": pts[0], "coordinate_y": pts[1]})
    for k, v in enumerate(vels):
        savemat(tmp_path / f"v{k:02d}.mat", {"velocity_x": v[0], "velocity_y": v[1]})
    x0 = H.seeds_2d(30, 15)
    n = x0.shape[0] // 4
    for j in range(2):
        blocks = {key: x0[i * n:(i + 1) * n] + 0.01 * j for i, key in enumerate(("left", "right", "top", "bottom"))}
        savemat(tmp_path / f"p{j}.mat", blocks)
    (tmp_path / "v.txt").write_text("\n".join(str(tmp_path / f"v{k:02d}.mat") for k in range(12)))
    (tmp_path / "c.txt").write_text(str(tmp_path / "c.mat"))
    (tmp_path / "p.txt").write_text("\n".join(str(tmp_path / f"p{j}.mat") for j in range(2)))
    assert cx.shape == (ny, nx) and cy.shape == (ny, nx)
    for dt in (0.25, -0.25):
        cfg = load_config(None, experiment_name="t", list_velocity_files=str(tmp_path / "v.txt"),
                          list_coordinate_files=str(tmp_path / "c.txt"),
                          list_particle_files=str(tmp_path / "p.txt"), snapshot_timestep=dt,
                          flow_map_period=1.0, integrator=method, interpolator="grid",
                          flow_grid_shape=f"{nx},{ny}")
        m = FileFTLEManager(cfg, write=False)
        batched = m.run(batched=True)
        single = m.run(batched=False)
        assert len(batched) == len(single) == 12 - 5 + 1
        for a, b in zip(batched, single):
            np.testing.assert_array_equal(a, b)


def test_analytical_solver_gridded_field_converges_to_exact_oracle():
    """AnalyticalSolver (api.py:15-224): the user's function, frozen at time 0
    (interpolate.py:788,812), is sampled onto a device grid; the FTLE then
    differs from the oracle's exact function evaluation only by the bilinear
    interpolation error, O(spacing^2): within 1e-3 of the field range on a
    1024x512 grid, and about 4x smaller when the spacing halves."""
    from oracle import flows
    from pyftle_amd import NeighboringParticles
    from pyftle_amd.api import AnalyticalSolver

    x0 = H.seeds_2d(40, 20)
    fn = lambda x: flows.double_gyre(0.0, x)  # noqa: E731
    exp_ftle, _ = O.run_window_reference([fn] * 11, x0, 0.1, "rk4")
    rng = np.nanmax(exp_ftle) - np.nanmin(exp_ftle)
    errs = []
    for shape in ((1024, 512), (2048, 1024)):
        s = AnalyticalSolver(flows.double_gyre, NeighboringParticles(x0.copy()), 0.1, 1.0, 1, "rk4",
                             grid=((-0.5, -0.25), (2.5, 1.25), shape), precision="float64")
        got = s.run()
        assert got.shape == exp_ftle.shape
        errs.append(np.max(np.abs(got - exp_ftle)) / rng)
    assert errs[0] <= 1e-3
    assert errs[1] <= 0.35 * errs[0]


@pytest.mark.parametrize("method", ["rk4", "euler", "ab2"])
def test_analytical_solver_default_evaluates_function_at_every_stage(method):
    """AnalyticalSolver (api.py:15-224) with its defaults: velocity_fn(0.0, points) on the
    device stage positions at every stage (interpolate.py:788,812; integrate.py:291-342),
    against the oracle's exact-function window (run_window_reference with the NumPy
    double gyre).  FP64; the two sin/cos implementations differ by ulps only."""
    from oracle import flows
    from pyftle_amd import NeighboringParticles
    from pyftle_amd.api import AnalyticalSolver

    x0 = H.seeds_2d(40, 20)
    fn = lambda x: flows.double_gyre(0.0, x)  # noqa: E731
    exp_ftle, exp_pos = O.run_window_reference([fn] * 11, x0, 0.1, method)
    parts = NeighboringParticles(x0.copy())
    s = AnalyticalSolver(H.torch_double_gyre, parts, 0.1, 1.0, 1, method)
    got = s.run()
    rng = np.nanmax(exp_ftle) - np.nanmin(exp_ftle)
    err = np.max(np.abs(got - exp_ftle)) / rng
    print(f"[analytical {method}] max|err|/range = {err:.3e}")
    assert err <= 1e-10


def test_analytical_solver_windows_and_backward_helical():
    """The reference's helical-vortex example (examples/01_Closed-form_velocity_helical.ipynb),
    scaled down to 8x8x16 seeds: backward time (dt = -0.01, T = 0.02 -> 3 steps), Euler as in
    the notebook and RK4, two windows.  The function normalises by the largest speed over the
    queried points, so it must see the stage positions exactly as in the reference."""
    from oracle import flows
    from pyftle_amd import NeighboringParticles
    from pyftle_amd.api import AnalyticalSolver

    x0 = H.helical_particles(8, 8, 16)
    fn = lambda x: flows.helical_vortex(0.0, x)  # noqa: E731
    for method in ("euler", "rk4"):
        exp_ftle, _ = O.run_window_reference([fn] * 3, x0, -0.01, method)
        s = AnalyticalSolver(H.torch_helical_vortex, NeighboringParticles(x0.copy()), -0.01, 0.02, 2, method)
        got = s.run()
        assert got.shape == (2, exp_ftle.shape[0])
        rng = np.nanmax(exp_ftle) - np.nanmin(exp_ftle)
        for w in range(2):  # every window starts from the original particles at frozen time 0
            err = np.max(np.abs(got[w] - exp_ftle)) / rng
            print(f"[helical {method} window {w}] max|err|/range = {err:.3e}")
            assert err <= 1e-10


def test_analytical_integrator_step_and_errors():
    """Integrator.integrate(h, particles) on host positions with the analytical strategy
    equals one oracle step; a NumPy-only function raises TypeError on the device path; the
    opt-in grid raises when a particle leaves the sampled box."""
    from oracle import flows
    from pyftle_amd import NeighboringParticles
    from pyftle_amd.api import AnalyticalSolver
    from pyftle_amd.integrate import create_integrator
    from pyftle_amd.interpolate import create_interpolator

    x0 = H.seeds_2d(6, 4)
    integ = create_integrator("rk4", create_interpolator("analytical", velocity_fn=H.torch_double_gyre))
    parts = NeighboringParticles(x0.copy())
    integ.integrate(0.1, parts)
    exp = x0.copy()
    O.rk4_step(exp, 0.1, lambda x: flows.double_gyre(0.0, x))
    np.testing.assert_allclose(parts.positions, exp, rtol=0, atol=1e-14)
    # AnalyticalSolver.run reports a failed window as RuntimeError (parallel.py:236-269)
    with pytest.raises(RuntimeError) as ei:
        AnalyticalSolver(flows.double_gyre, NeighboringParticles(x0.copy()), 0.1, 0.2, 1, "rk4").run()
    assert isinstance(ei.value.__cause__, TypeError) and "torch" in str(ei.value.__cause__)
    far = NeighboringParticles(x0.copy())
    s = AnalyticalSolver(lambda t, q: 0 * q + 1.0, far, 0.1, 1.0, 1, "euler",
                         grid=((0.0, 0.0), (2.0, 1.0), (64, 32)), precision="float64")
    with pytest.raises(RuntimeError) as ei:
        s.run()
    assert isinstance(ei.value.__cause__, ValueError) and "left the sampled box" in str(ei.value.__cause__)
