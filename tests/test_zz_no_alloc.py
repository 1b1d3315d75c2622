"""GPU: the allocation-free batch entry point (pyftle_batch_run_noalloc) against
pyftle_batch_create + pyftle_batch_launch on the same windows, bit for bit, on two
calls in a row (the second smaller), through the C ABI."""

import ctypes

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyftle_amd import _native

    _native.load()


def _windows(specs, layout):
    from pyftle_amd.batch import Window
    from pyftle_amd.field import DeviceField, GridGeometry

    wins = []
    for grid, n_snap, (sx, sy), steps, h, t0, eps in specs:
        vels, pts = H.double_gyre_window(nx=grid[0], ny=grid[1], n_snap=n_snap, eps=eps)
        f = DeviceField(GridGeometry.from_reference(grid, pts), dtype="float32", layout=layout,
                        ring_slots=n_snap, n_snapshots=n_snap, t0=0.0, dt_snap=0.1)
        for k, v in enumerate(vels):
            f.pack(k, v)
        pos = torch.as_tensor(H.seeds_2d(sx, sy), device="cuda")
        wins.append(Window(field=f, positions=pos, n_steps=steps, h=h, t_start=t0, map_period=steps * abs(h),
                           positions_out=torch.empty_like(pos)))
    return wins


SPECS = [  # grid, n_snap, seeds (nsx, nsy), steps, h, t_start, eps
    ((32, 16), 9, (7, 3), 12, 0.05, 0.0, 0.25),
    ((64, 32), 13, (16, 8), 20, 0.05, 0.1, 0.1),
    ((48, 24), 21, (33, 5), 35, 0.05, 0.2, 0.3),
    ((40, 20), 6, (1, 1), 8, 0.05, 0.0, 0.2),
    ((96, 48), 11, (50, 25), 18, -0.05, 1.0, 0.15),
    ((128, 64), 16, (64, 40), 28, 0.05, 0.05, 0.22),
]
SMALL = [
    ((48, 24), 9, (5, 4), 10, 0.05, 0.1, 0.2),
    ((32, 16), 7, (3, 2), 6, -0.05, 0.6, 0.25),
]


def _fresh_outputs(wins):
    for w in wins:
        w.ftle = torch.full((w.positions.shape[0] // 4,), -7.0, dtype=torch.float64, device="cuda")
        w.positions_out = torch.full_like(w.positions, -7.0)


@pytest.mark.parametrize("layout_name", ["LAYOUT_PAIR", "LAYOUT_CELL_PAIR"])
def test_noalloc_batch_matches_batch_plan_twice(layout_name):
    from pyftle_amd import _native as N
    from pyftle_amd.batch import BatchConfig, BatchPlan, run_batch_noalloc

    layout = getattr(N, layout_name)
    for call, specs in enumerate((SPECS, SMALL)):
        wins = _windows(specs, layout)
        cfg = BatchConfig(groups_per_block=64, blocks_per_cu=2, use_graph=0)
        plan = BatchPlan(wins, method="rk4", time_mode="linear", cfg=cfg)
        plan.run()
        torch.cuda.synchronize()
        exp = [(w.ftle.clone(), w.positions_out.clone()) for w in wins]
        plan.close()
        _fresh_outputs(wins)
        print(f"[gpu] noalloc call {call + 1}: {len(wins)} windows, {layout_name}", flush=True)
        got = run_batch_noalloc(wins, method="rk4", time_mode="linear", cfg=cfg)
        torch.cuda.synchronize()
        for i, (w, g, (ef, ep)) in enumerate(zip(wins, got, exp)):
            assert torch.equal(g, ef), f"call {call + 1} window {i}: FTLE differs"
            assert torch.equal(w.positions_out, ep), f"call {call + 1} window {i}: positions differ"
            assert np.isfinite(g.cpu().numpy()).all()


def test_noalloc_batch_rejects_too_many_fields():
    from pyftle_amd import _native as N
    from pyftle_amd.batch import _SIGS, NOALLOC_MAX_FIELDS, BatchConfig, _describe, run_batch_noalloc

    N.extra_signatures(_SIGS)
    wins = _windows(SMALL[:1], N.LAYOUT_PAIR)
    arr, o = _describe(wins, "rk4", "linear", BatchConfig())
    # the count is checked before any descriptor is read
    rc = N.load().pyftle_batch_run_noalloc(arr, NOALLOC_MAX_FIELDS + 1, ctypes.byref(o), None)
    assert rc == N.EINVAL
    assert b"at most" in N.load().pyftle_last_error()
    with pytest.raises(ValueError):
        run_batch_noalloc(wins * (NOALLOC_MAX_FIELDS + 1))
