"""GPU: the scattered-data strategies (interpolate.py:114-294) against the oracle, which
makes the reference's SciPy calls (LinearNDInterpolator on the Delaunay triangulation,
NearestNDInterpolator, CloughTocher2DInterpolator).

* sampling in 2D and 3D, FP64 and FP32 records: values, NaN outside the convex hull
  (linear), defined everywhere (nearest);
* a whole FTLESolver window on scattered points (FP64), RK4 and Euler, against the
  oracle's window driver on the same SciPy interpolators.
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


def _scattered_2d(n=1500, seed=0, t=0.3):
    rng = np.random.default_rng(seed)
    pts = np.vstack([rng.uniform(0.0, 2.0, n), rng.uniform(0.0, 1.0, n)])
    corners = np.array([[0.0, 2.0, 0.0, 2.0], [0.0, 0.0, 1.0, 1.0]])
    pts = np.hstack([corners, pts])  # the hull is the whole domain
    vel = flows.double_gyre(t, pts.T).T.copy()
    return vel, pts


def _scattered_3d(n=900, seed=1):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0.0, 2.0 * np.pi, (3, n))
    vel = flows.abc(0.0, pts.T).T.copy()
    return vel, pts


@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("method,dim", [("linear", 2), ("linear", 3), ("nearest", 2), ("nearest", 3),
                                        ("cubic", 2)])
def test_scattered_sampling_matches_scipy(dim, method, precision):
    from pyftle_amd.interpolate import create_interpolator

    vel, pts = _scattered_2d() if dim == 2 else _scattered_3d()
    g = create_interpolator(method, precision=precision)
    g.update(vel, pts)
    rng = np.random.default_rng(5)
    lo, hi = pts.min(axis=1), pts.max(axis=1)
    q = rng.uniform(lo - 0.2 * (hi - lo), hi + 0.2 * (hi - lo), (5000, dim))
    q = np.vstack([q, pts.T[:40]])  # data points themselves
    got = g.interpolate(q)
    exp = O.ScatteredField(pts, vel, method).interpolate(q)
    nan_g, nan_e = np.isnan(got).any(axis=1), np.isnan(exp).any(axis=1)
    scale = np.nanmax(np.abs(exp))
    ok = ~nan_e & ~nan_g
    err = np.abs(got[ok] - exp[ok]).max() / scale
    exact = np.mean(np.all(got[ok] == exp[ok], axis=1))
    print(f"[gpu] scattered {dim}D {method} {precision}: max err {err:.2e}, bit-equal rows {exact:.4f}, "
          f"NaN rows gpu/scipy {nan_g.sum()}/{nan_e.sum()}", flush=True)
    if method == "nearest":
        assert not nan_g.any() and not nan_e.any()
    elif method == "cubic" and precision == "float32":
        assert np.array_equal(nan_g, nan_e)
    else:
        assert nan_e.sum() > 100  # queries outside the hull exist
        assert np.array_equal(nan_g, nan_e)
    if precision == "float64":
        tol = 1e-13 if method == "cubic" else 1e-14
    else:
        tol = 1e-6 if method == "cubic" else 1e-7
    assert err <= tol


@pytest.mark.parametrize("method,integrator", [("linear", "rk4"), ("nearest", "rk4"), ("linear", "euler"),
                                               ("cubic", "rk4")])
def test_ftle_solver_window_on_scattered_points(method, integrator):
    from pyftle_amd import NeighboringParticles
    from pyftle_amd.data_source import ArrayBatchSource
    from pyftle_amd.ftle_solver import FTLESolver
    from pyftle_amd.integrate import create_integrator
    from pyftle_amd.interpolate import create_interpolator

    p = 7
    snaps = [_scattered_2d(n=2500, seed=3, t=0.1 * k) for k in range(p)]
    pts = snaps[0][1]
    vels = [v for v, _ in snaps]
    x0 = H.seeds_2d(12, 6)
    fields = [O.ScatteredField(pts, v, method).interpolate for v in vels]
    exp_ftle, exp_pos = O.run_window_reference(fields, x0, 0.1, integrator)
    interp = create_interpolator(method, precision="float64")
    parts = NeighboringParticles(x0.copy())
    ftle = FTLESolver(ArrayBatchSource(vels, pts, parts, 0.1), create_integrator(integrator, interp)).run()
    pos_err = np.abs(parts.positions - exp_pos).max()
    ok = np.isfinite(exp_ftle)
    f_err = np.abs(ftle[ok] - exp_ftle[ok]).max() / (exp_ftle[ok].max() - exp_ftle[ok].min())
    print(f"[gpu] scattered window {method}/{integrator}: positions {pos_err:.2e}, FTLE {f_err:.2e}", flush=True)
    assert pos_err <= 1e-12
    assert f_err <= 1e-8
