"""GPU: GridInterpolator (the reference's SciPy RegularGridInterpolator strategies,
interpolate.py:546-762) against the oracle, which makes the reference's SciPy calls.

* sampling, 2D and 3D, FP64 and FP32 records: ``linear`` and ``nearest`` against
  ``RegularGridInterpolator`` itself; ``cubic`` against the exact not-a-knot spline
  (``solver=spsolve``) and, looser, against SciPy's default iterative fit, which the
  reference uses;
* outside the grid -> 0, the upper edge inside, NaN -> NaN;
* a whole FTLESolver window (RK4/Euler/AB2, FP64) against the oracle's window
  driver running on the same SciPy interpolators;
* a file experiment with ``interpolator: cubic`` and ``flow_grid_shape``, the
  reference's sample configuration.
"""

import numpy as np
import pytest

from oracle import flows
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


def _gyre_c_order(nx, ny, t):
    """Double gyre on a C-ordered (nx, ny) grid over [0, 2] x [0, 1]: points and
    velocities as (2, nx*ny) arrays, index ix*ny + iy."""
    x, y = np.linspace(0.0, 2.0, nx), np.linspace(0.0, 1.0, ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel()])
    vel = flows.double_gyre(t, pts.T).T.copy()
    return vel, pts


def _abc_c_order(n):
    ax = np.linspace(0.2, 2.0 * np.pi - 0.3, n[0]), np.linspace(0.1, 6.0, n[1]), np.linspace(0.0, 5.5, n[2])
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    vel = flows.abc(0.0, pts.T).T.copy()
    return vel, pts


def _queries(lo, hi, pts, n=4000, seed=0):
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo), np.asarray(hi)
    q = rng.uniform(lo, hi, (n, len(lo)))
    nodes = pts.T[rng.integers(0, pts.shape[1], 50)]
    outside = np.vstack([lo - 0.1, hi + 0.1, [lo[0] - 1e-9] + list(lo[1:])])
    return np.vstack([q, nodes, hi[None, :], lo[None, :], outside])


@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("method", ["linear", "nearest", "cubic"])
def test_sampling_matches_scipy_2d(method, precision):
    from scipy.sparse.linalg import spsolve

    from pyftle_amd.interpolate import create_interpolator

    nx, ny = 40, 24
    vel, pts = _gyre_c_order(nx, ny, 0.3)
    g = create_interpolator(method, (nx, ny), precision=precision)
    g.update(vel, pts)
    q = _queries([0.0, 0.0], [2.0, 1.0], pts)
    got = g.interpolate(q)
    exact = O.RGIField((nx, ny), vel, pts, method, solver=spsolve if method == "cubic" else None).interpolate(q)
    scale = np.abs(exact).max()
    err = np.abs(got - exact).max() / scale
    print(f"[gpu] 2D {method} {precision}: max |gpu - scipy| / scale = {err:.2e}", flush=True)
    if precision == "float64":
        tol = {"linear": 1e-15, "nearest": 0.0, "cubic": 1e-13}[method]
    else:
        tol = {"linear": 2e-7, "nearest": 1e-7, "cubic": 1e-6}[method]
    assert err <= tol
    assert np.all(got[-3:] == 0.0)  # outside on either side
    if method == "cubic":
        ref = O.RGIField((nx, ny), vel, pts, "cubic").interpolate(q)  # the reference's default solver
        gap = np.abs(ref - exact).max() / scale
        print(f"[gpu] 2D cubic: SciPy default (iterative) fit vs exact spline {gap:.2e}", flush=True)
        assert np.abs(got - ref).max() / scale <= 1e-3
    nan = g.interpolate(np.array([[np.nan, 0.5], [1.0, 0.5]]))
    assert np.isnan(nan[0]).all() and np.isfinite(nan[1]).all()


@pytest.mark.parametrize("method", ["linear", "nearest", "cubic"])
def test_sampling_matches_scipy_3d(method):
    from scipy.sparse.linalg import spsolve

    from pyftle_amd.interpolate import GridInterpolator

    n = (14, 11, 9)
    vel, pts = _abc_c_order(n)
    g = GridInterpolator(n, method, precision="float64")
    g.update(vel, pts)
    lo, hi = pts.min(axis=1), pts.max(axis=1)
    q = _queries(lo, hi, pts, n=3000, seed=1)
    got = g.interpolate(q)
    exact = O.RGIField(n, vel, pts, method, solver=spsolve if method == "cubic" else None).interpolate(q)
    err = np.abs(got - exact).max() / np.abs(exact).max()
    print(f"[gpu] 3D {method}: max |gpu - scipy| / scale = {err:.2e}", flush=True)
    assert err <= {"linear": 1e-15, "nearest": 0.0, "cubic": 1e-13}[method]
    assert np.all(got[-3:] == 0.0)


@pytest.mark.parametrize("method,integrator", [("linear", "rk4"), ("cubic", "rk4"), ("nearest", "rk4"),
                                               ("cubic", "euler"), ("linear", "ab2")])
def test_ftle_solver_window_with_grid_interpolator(method, integrator):
    """FTLESolver + GridInterpolator (FP64) == the oracle's window driver on SciPy's
    RegularGridInterpolator, snapshot i frozen for step i (ftle_solver.py:99-102)."""
    from scipy.sparse.linalg import spsolve

    from pyftle_amd import NeighboringParticles
    from pyftle_amd.data_source import ArrayBatchSource
    from pyftle_amd.ftle_solver import FTLESolver
    from pyftle_amd.integrate import create_integrator
    from pyftle_amd.interpolate import create_interpolator

    nx, ny, p = 48, 24, 8
    snaps = [_gyre_c_order(nx, ny, 0.1 * k) for k in range(p)]
    pts = snaps[0][1]
    vels = [v for v, _ in snaps]
    x0 = H.seeds_2d(14, 7)
    solver = spsolve if method == "cubic" else None
    fields = [O.RGIField((nx, ny), v, pts, method, solver=solver).interpolate for v in vels]
    exp_ftle, exp_pos = O.run_window_reference(fields, x0, 0.1, integrator)
    interp = create_interpolator(method, (nx, ny), precision="float64")
    parts = NeighboringParticles(x0.copy())
    ftle = FTLESolver(ArrayBatchSource(vels, pts, parts, 0.1), create_integrator(integrator, interp)).run()
    pos_err = np.abs(parts.positions - exp_pos).max()
    ok = np.isfinite(exp_ftle)
    f_err = np.abs(ftle[ok] - exp_ftle[ok]).max() / (exp_ftle[ok].max() - exp_ftle[ok].min())
    print(f"[gpu] window {method}/{integrator}: positions {pos_err:.2e}, FTLE {f_err:.2e} of range", flush=True)
    assert pos_err <= 1e-12
    assert f_err <= 1e-8


def test_file_experiment_with_cubic_grid_interpolator(tmp_path):
    """The reference's sample config (interpolator: cubic with flow_grid_shape) runs:
    each window through FTLESolver on the cubic mode, equal to the oracle's windows."""
    from scipy.io import savemat
    from scipy.sparse.linalg import spsolve

    from pyftle_amd.experiment import FileFTLEManager, load_config

    nx, ny = 32, 16
    vf = []
    vels = []
    for k in range(6):
        v, pts = _gyre_c_order(nx, ny, 0.25 * k)
        vels.append(v)
        savemat(tmp_path / f"v{k}.mat", {"velocity_x": v[0].reshape(nx, ny), "velocity_y": v[1].reshape(nx, ny)})
        vf.append(str(tmp_path / f"v{k}.mat"))
    savemat(tmp_path / "c.mat", {"coordinate_x": pts[0].reshape(nx, ny), "coordinate_y": pts[1].reshape(nx, ny)})
    x0 = H.seeds_2d(10, 5)
    n = x0.shape[0] // 4
    savemat(tmp_path / "p.mat", {"left": x0[:n], "right": x0[n:2 * n], "top": x0[2 * n:3 * n], "bottom": x0[3 * n:]})
    for name, items in (("v.txt", vf), ("c.txt", [str(tmp_path / "c.mat")] * 6), ("p.txt", [str(tmp_path / "p.mat")])):
        (tmp_path / name).write_text("\n".join(items))
    cfg = load_config(None, experiment_name="t", list_velocity_files=str(tmp_path / "v.txt"),
                      list_coordinate_files=str(tmp_path / "c.txt"), list_particle_files=str(tmp_path / "p.txt"),
                      snapshot_timestep=0.25, flow_map_period=0.75, integrator="rk4", interpolator="cubic",
                      flow_grid_shape=f"{nx},{ny}")
    m = FileFTLEManager(cfg, precision="float64", write=False)
    res = m.run()
    p = m.window_length()
    assert len(res) == 6 - p + 1
    for i, r in enumerate(res):
        fields = [O.RGIField((nx, ny), vels[i + k], pts, "cubic", solver=spsolve).interpolate for k in range(p)]
        exp, _ = O.run_window_reference(fields, x0, 0.25, "rk4")
        np.testing.assert_allclose(r, exp, rtol=1e-9, atol=1e-12)
