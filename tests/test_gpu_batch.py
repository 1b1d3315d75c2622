"""GPU: batched persistent launch == one launch per window, graph replays stay exact."""

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


def _pair_field(vels, pts, grid_shape, dt_snap=0.1):
    from pyftle_amd import _native as N
    from pyftle_amd.field import DeviceField, GridGeometry

    f = DeviceField(GridGeometry.from_reference(grid_shape, pts), dtype="float32", layout=N.LAYOUT_PAIR,
                    ring_slots=len(vels), n_snapshots=len(vels), t0=0.0, dt_snap=dt_snap)
    for k, v in enumerate(vels):
        f.pack(k, v)
    return f


def _windows_2d():
    from pyftle_amd.batch import Window

    specs = [  # grid, n_snap, seeds (nsx, nsy), steps, h, t_start, eps
        ((32, 16), 9, (7, 3), 12, 0.05, 0.0, 0.25),
        ((64, 32), 13, (16, 8), 20, 0.05, 0.1, 0.1),
        ((48, 24), 21, (33, 5), 35, 0.05, 0.2, 0.3),
        ((40, 20), 6, (1, 1), 8, 0.05, 0.0, 0.2),
        ((96, 48), 11, (50, 25), 18, -0.05, 1.0, 0.15),
    ]
    wins, host = [], []
    for grid, n_snap, (sx, sy), steps, h, t0, eps in specs:
        vels, pts = H.double_gyre_window(nx=grid[0], ny=grid[1], n_snap=n_snap, eps=eps)
        x0 = H.seeds_2d(sx, sy)
        f = _pair_field(vels, pts, grid)
        pos = torch.as_tensor(x0, device="cuda")
        wins.append(Window(field=f, positions=pos, n_steps=steps, h=h, t_start=t0, map_period=steps * abs(h)))
        host.append((grid, vels, pts, x0, steps, h, t0))
    return wins, host


@pytest.mark.parametrize("use_graph", [1, 0])
def test_batch_matches_single_window_launches(use_graph):
    from pyftle_amd import window as W
    from pyftle_amd.batch import BatchConfig, BatchPlan

    wins, _ = _windows_2d()
    for g in (32, 64):
        plan = BatchPlan(wins, method="rk4", time_mode="linear",
                         cfg=BatchConfig(groups_per_block=g, use_graph=use_graph, blocks_per_cu=2))
        print(f"[gpu] batch G={g} graph={use_graph} {plan.info()}", flush=True)
        plan.run()
        for w in wins:
            single = W.advect_ftle(w.field, w.positions, method="rk4", time_mode="linear", n_steps=w.n_steps,
                                   h=w.h, t_start=w.t_start, map_period=w.map_period,
                                   cfg=W.LaunchConfig(groups_per_block=g))
            assert torch.equal(single, w.ftle)
        plan.close()


def test_batch_graph_replays_are_exact():
    from pyftle_amd.batch import BatchConfig, BatchPlan

    wins, _ = _windows_2d()
    plan = BatchPlan(wins, cfg=BatchConfig(use_graph=1))
    plan.run()
    torch.cuda.synchronize()
    eager = [w.ftle.clone() for w in wins]
    for _ in range(4):
        for w in wins:
            w.ftle.fill_(-1.0)  # eager work between replays; stale output would show
        torch.cuda.synchronize()
        plan.run()
        torch.cuda.synchronize()
        for w, e in zip(wins, eager):
            assert torch.equal(w.ftle, e)
    plan.close()


def test_batch_against_oracle():
    from pyftle_amd.batch import BatchPlan

    wins, host = _windows_2d()
    plan = BatchPlan(wins)
    plan.run()
    for w, (grid, vels, pts, x0, steps, h, t0) in zip(wins, host):
        fields = H.grid_fields(grid, vels, pts)
        vat = O.time_linear_velocity([f.interpolate for f in fields], 0.0, 0.1)
        exp, _ = O.run_window_linear(vat, x0, t0, h, steps, "rk4")
        got = w.ftle.cpu().numpy()
        # a one-seed window has zero range: scale by the field magnitude there
        scale = max(exp.max() - exp.min(), np.abs(exp).max())
        assert np.max(np.abs(got - exp)) / scale <= 1e-3
    plan.close()


def test_batch_3d_steady_periodic():
    from pyftle_amd import _native as N
    from pyftle_amd import window as W
    from pyftle_amd.batch import BatchPlan, Window
    from pyftle_amd.field import DeviceField, GridGeometry

    wins = []
    for n, ns, steps in ((16, 3, 30), (24, 4, 45)):
        vels, pts = H.abc_window(n=n)
        f = DeviceField(GridGeometry.from_reference((n, n, n), pts, periodic=(True, True, True)),
                        dtype="float32", layout=N.LAYOUT_SINGLE)
        f.pack(0, vels[0])
        pos = torch.as_tensor(H.seeds_3d(ns), device="cuda")
        wins.append(Window(field=f, positions=pos, n_steps=steps, h=0.05, map_period=steps * 0.05))
    plan = BatchPlan(wins, method="rk4", time_mode="steady")
    plan.run()
    for w in wins:
        single = W.advect_ftle(w.field, w.positions, method="rk4", time_mode="steady", n_steps=w.n_steps,
                               h=w.h, map_period=w.map_period)
        assert torch.equal(single, w.ftle)
    plan.close()


def test_batch_rejects_mixed_dims():
    from pyftle_amd import _native as N
    from pyftle_amd.batch import BatchPlan, Window
    from pyftle_amd.field import DeviceField, GridGeometry

    wins, _ = _windows_2d()
    vels, pts = H.abc_window(n=8)
    f3 = DeviceField(GridGeometry.from_reference((8, 8, 8), pts), dtype="float32", layout=N.LAYOUT_PAIR)
    wins.append(Window(field=f3, positions=torch.as_tensor(H.seeds_3d(2), device="cuda"), n_steps=3,
                       h=0.05, map_period=0.15))
    with pytest.raises(RuntimeError, match="share dim"):
        BatchPlan(wins)


def test_batch_rejects_malformed_windows():
    """Shapes that do not match the block layout are refused before any launch."""
    from pyftle_amd.batch import BatchPlan, Window

    wins, _ = _windows_2d()
    w = wins[0]
    bad = [
        Window(field=w.field, positions=w.positions[:-1].contiguous(), n_steps=4, h=0.05, map_period=0.2),
        Window(field=w.field, positions=w.positions.reshape(-1), n_steps=4, h=0.05, map_period=0.2),
        Window(field=w.field, positions=w.positions, n_steps=4, h=0.05, map_period=0.2,
               positions_out=torch.empty(3, 2, dtype=w.positions.dtype, device="cuda")),
        Window(field=w.field, positions=w.positions, n_steps=4, h=0.05, map_period=0.2,
               ftle=torch.empty(5, dtype=torch.float32, device="cuda")),
    ]
    for b in bad:
        with pytest.raises(ValueError):
            BatchPlan([b], method="rk4", time_mode="linear")
    with pytest.raises(ValueError):
        BatchPlan([], method="rk4")


def test_batch_counting_launch_equals_single_window_counts():
    """The counting copy of the persistent kernel (pyftle_batch_count_loads) leaves every
    window's FTLE bit-identical, and the record loads it counts equal the sum of the counts of
    the single-window kernel over the same windows at the same block shape (a work item is
    one block of one window, so its waves are the single-window kernel's waves)."""
    from pyftle_amd import _native as N
    from pyftle_amd import window as W
    from pyftle_amd.batch import BatchConfig, BatchPlan, Window
    from pyftle_amd.field import DeviceField, GridGeometry

    wins = []
    for (nx, ny), n_snap, (sx, sy), steps, t0 in (((64, 32), 13, (40, 20), 20, 0.1), ((48, 24), 21, (33, 5), 35, 0.2),
                                                  ((96, 48), 11, (50, 25), 16, 0.0)):
        vels, pts = H.double_gyre_window(nx=nx, ny=ny, n_snap=n_snap)
        f = DeviceField(GridGeometry.from_reference((nx, ny), pts), dtype="float32", layout=N.LAYOUT_CELL_PAIR,
                        ring_slots=n_snap, n_snapshots=n_snap, t0=0.0, dt_snap=0.1)
        for k, v in enumerate(vels):
            f.pack(k, v)
        pos = torch.as_tensor(H.seeds_2d(sx, sy), device="cuda")
        wins.append(Window(field=f, positions=pos, n_steps=steps, h=0.05, t_start=t0, map_period=steps * 0.05))
    plan = BatchPlan(wins, method="rk4", time_mode="linear", cfg=BatchConfig(groups_per_block=32, use_graph=1))
    plan.run()
    torch.cuda.synchronize()
    before = [w.ftle.clone() for w in wins]
    loads = plan.count_loads()
    assert all(torch.equal(b, w.ftle) for b, w in zip(before, wins))
    total = 0
    for w in wins:
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        single = W.advect_ftle(w.field, w.positions, method="rk4", time_mode="linear", n_steps=w.n_steps, h=w.h,
                               t_start=w.t_start, map_period=w.map_period, count_loads=cnt,
                               cfg=W.LaunchConfig(groups_per_block=32, xcd_remap=0))
        assert torch.equal(single, w.ftle)
        total += int(cnt.item())
    print(f"[gpu] batch counted {loads} record wave loads, single windows {total}", flush=True)
    assert loads == total > 0
    plan.close()
