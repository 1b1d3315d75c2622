"""GPU edge cases of the window kernel against the FP64 oracle.

* seeds outside the grid and exactly on its upper boundary: zero velocity there
  (GridInterp.cpp:44-58,116-126), so those particles stay put;
* backward time (negative h, the reference's negative snapshot_timestep);
* ragged seed counts (1, 63, 65, 257 seeds: partial work-groups);
* empty windows (no seeds) and zero-step windows (identity map, FTLE 0).

FP64 fields: positions bit-exact, FTLE to 1e-12; FP32: FTLE within 1e-3 of the range.
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


def _field(vels, pts, shape, precision):
    from pyftle_amd import _native as N
    from pyftle_amd.field import DeviceField, GridGeometry

    f = DeviceField(GridGeometry.from_reference(shape, pts), dtype=precision, layout=N.LAYOUT_SINGLE,
                    ring_slots=len(vels), n_snapshots=len(vels))
    for k, v in enumerate(vels):
        f.pack(k, v)
    return f


def _groups(centers, delta=1e-3):
    return O.make_neighbor_positions(np.asarray(centers, dtype=np.float64), delta)


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_outside_and_upper_boundary_seeds(precision):
    from pyftle_amd import window as W

    vels, pts = H.double_gyre_window(nx=48, ny=24, n_snap=6)
    # domain [0, 2] x [0, 1]: inside, outside on every side, exactly on the upper edges
    centers = [[1.0, 0.5], [-0.5, 0.5], [2.5, 0.5], [1.0, -0.3], [1.0, 1.4], [2.0, 0.5], [1.0, 1.0],
               [2.0, 1.0], [0.0, 0.0], [1.9995, 0.9995]]
    x0 = _groups(centers)
    fields = H.grid_fields((48, 24), vels, pts)
    exp_ftle, exp_pos = O.run_window_reference([f.interpolate for f in fields], x0, 0.1, "rk4")
    f = _field(vels, pts, (48, 24), precision)
    pos = torch.as_tensor(x0, device="cuda")
    out = torch.empty_like(pos)
    ftle = W.advect_ftle(f, pos, method="rk4", time_mode="frozen", n_steps=6, h=0.1, map_period=0.5,
                         positions_out=out)
    got = ftle.cpu().numpy()
    if precision == "float64":
        np.testing.assert_array_equal(out.cpu().numpy(), exp_pos)
        np.testing.assert_allclose(got, exp_ftle, rtol=1e-12, atol=1e-14)
    else:
        ok = np.isfinite(exp_ftle)
        rng = np.nanmax(exp_ftle) - np.nanmin(exp_ftle)
        assert np.max(np.abs(got[ok] - exp_ftle[ok])) <= 1e-3 * rng
    # particles fully outside never move (the FP32 path re-forms the position from its
    # cell-local coordinates: equal to ~1e-15 relative)
    P = out.cpu().numpy().reshape(4, -1, 2)
    X0 = x0.reshape(4, -1, 2)
    for s in (1, 2, 3, 4):
        if precision == "float64":
            np.testing.assert_array_equal(P[:, s], X0[:, s])
        else:
            np.testing.assert_allclose(P[:, s], X0[:, s], rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("method", ["rk4", "euler", "ab2"])
def test_backward_time_window(method):
    from pyftle_amd import window as W

    vels, pts = H.double_gyre_window(nx=64, ny=32, n_snap=9)
    vels = vels[::-1]  # backward FTLE: snapshots in reverse, negative step
    x0 = H.seeds_2d(12, 6)
    fields = H.grid_fields((64, 32), vels, pts)
    exp_ftle, exp_pos = O.run_window_reference([f.interpolate for f in fields], x0, -0.1, method)
    f = _field(vels, pts, (64, 32), "float64")
    pos = torch.as_tensor(x0, device="cuda")
    out = torch.empty_like(pos)
    ftle = W.advect_ftle(f, pos, method=method, time_mode="frozen", n_steps=9, h=-0.1, map_period=0.8,
                         positions_out=out)
    np.testing.assert_array_equal(out.cpu().numpy(), exp_pos)
    np.testing.assert_allclose(ftle.cpu().numpy(), exp_ftle, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("n_seeds", [1, 63, 65, 257])
def test_ragged_seed_counts(n_seeds):
    from pyftle_amd import window as W

    vels, pts = H.double_gyre_window(nx=64, ny=32, n_snap=5)
    rng = np.random.default_rng(n_seeds)
    centers = rng.uniform([0.1, 0.1], [1.9, 0.9], (n_seeds, 2))
    x0 = _groups(centers)
    fields = H.grid_fields((64, 32), vels, pts)
    exp_ftle, exp_pos = O.run_window_reference([f.interpolate for f in fields], x0, 0.1, "rk4")
    f = _field(vels, pts, (64, 32), "float64")
    out = torch.empty(x0.shape, dtype=torch.float64, device="cuda")
    ftle = W.advect_ftle(f, torch.as_tensor(x0, device="cuda"), method="rk4", time_mode="frozen", n_steps=5,
                         h=0.1, map_period=0.4, positions_out=out)
    assert ftle.shape == (n_seeds,)
    np.testing.assert_array_equal(out.cpu().numpy(), exp_pos)
    np.testing.assert_allclose(ftle.cpu().numpy(), exp_ftle, rtol=1e-12, atol=1e-14)


def test_empty_and_zero_step_windows():
    from pyftle_amd import window as W

    vels, pts = H.double_gyre_window(nx=32, ny=16, n_snap=2)
    f = _field(vels, pts, (32, 16), "float64")
    empty = W.advect_ftle(f, torch.zeros(0, 2, dtype=torch.float64, device="cuda"), method="rk4",
                          time_mode="frozen", n_steps=2, h=0.1, map_period=0.1)
    assert empty.shape == (0,)
    x0 = torch.as_tensor(H.seeds_2d(5, 3), device="cuda")
    out = torch.empty_like(x0)
    ftle = W.advect_ftle(f, x0, method="rk4", time_mode="frozen", n_steps=0, h=0.1, map_period=1.0,
                         positions_out=out)
    assert torch.equal(out, x0)
    assert torch.all(ftle == 0.0)


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_seed_order_leaves_results_unchanged(precision):
    """Processing the seed groups in Morton order (or any permutation) gives bit-identical
    FTLE and final positions: every group is independent and written at its own index."""
    from pyftle_amd import window as W

    vels, pts = H.double_gyre_window(nx=64, ny=32, n_snap=6)
    x0 = torch.as_tensor(H.seeds_2d(37, 19), device="cuda")
    f = _field(vels, pts, (64, 32), precision)
    kw = dict(method="rk4", time_mode="frozen", n_steps=6, h=0.1, map_period=0.5)
    out0, out1, out2 = (torch.empty_like(x0) for _ in range(3))
    base = W.advect_ftle(f, x0, positions_out=out0, **kw)
    morton = W.advect_ftle(f, x0, positions_out=out1, order="morton", **kw)
    n = x0.shape[0] // 4
    perm = torch.randperm(n, generator=torch.Generator().manual_seed(1)).to(torch.int32).cuda()
    rand = W.advect_ftle(f, x0, positions_out=out2, order=perm, **kw)
    order = W.seed_order(f, x0, 4)
    assert sorted(order.cpu().tolist()) == list(range(n))
    assert not torch.equal(order.cpu(), torch.arange(n, dtype=torch.int32))  # it does reorder
    for got, pos in ((morton, out1), (rand, out2)):
        assert torch.equal(got, base)
        assert torch.equal(pos, out0)
