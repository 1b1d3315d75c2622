"""GPU: the benchmark configurations at their full sizes.

* gyre512 -- 2D double gyre on 512x256, 101 snapshots t in [0, 10], 400x200 seeds,
  RK4 h = 0.05 over T = 10 in time-linear mode, FP32 cell-pair records;
* abc128 -- 3D steady ABC on a periodic 128^3 grid, 128^3 seeds, RK4 h = 0.01, 1000 steps;
* batch64 -- the 64-field batch of tests/golden/batch64_table.json in one launch.

Each FP32 field is checked against the FP64 kernel over every seed (max |err| <= 1e-3 of
the FP64 field's range) and the FP64 kernel against the NumPy oracle on sampled seeds
(positions bit for bit: same operation order, no contraction).  The batch is bit-equal to
single-window launches.
"""

import json
import os

import numpy as np
import pytest

from oracle import flows
from oracle import numpy_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "batch64_table.json")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyftle_amd import _native

    _native.load()


def _fp64_twin(wl_field, host_snaps):
    """The same time levels in FP64 node records (the kernel that matches the oracle)."""
    from pyftle_amd import _native as N
    from pyftle_amd.field import DeviceField

    pair = wl_field.layout in (N.LAYOUT_PAIR, N.LAYOUT_CELL_PAIR)
    f64 = DeviceField(wl_field.geom, dtype="float64", layout=N.LAYOUT_PAIR if pair else N.LAYOUT_SINGLE,
                      ring_slots=wl_field.ring_slots, n_snapshots=wl_field.n_snapshots, t0=wl_field.t0,
                      dt_snap=wl_field.dt_snap)
    for k, v in enumerate(host_snaps):
        f64.pack(k, v)
    return f64


def _sample(pos, n, nb, idx):
    return np.concatenate([pos[j * n + idx] for j in range(nb)], axis=0)


def _gyre_oracle_fields(snaps, nx, ny):
    fields = []
    for v in snaps:
        gf = O.GridField.__new__(O.GridField)
        gf.dim, gf.mins, gf.periodic = 2, [0.0, 0.0], (False, False)
        gf.v = [np.ascontiguousarray(v[c].reshape(ny, nx).T) for c in range(2)]
        gf.deltas = [2.0 / (nx - 1), 1.0 / (ny - 1)]
        fields.append(gf.interpolate)
    return fields


def _range_err(a, b):
    a, b = a.double(), b.double()
    return ((a - b).abs().max() / (b.max() - b.min())).item()


def test_gyre512_full_window_fp32_fp64_oracle():
    import bench
    from pyftle_amd import window as W

    wl = bench.build_gyre(torch)
    kw = dict(method="rk4", time_mode="linear", n_steps=wl["steps"], h=wl["h"], t_start=0.0,
              map_period=wl["map_period"])
    f32 = W.advect_ftle(wl["field"], wl["pos"], **kw)
    f64field = _fp64_twin(wl["field"], wl["host"]["snaps"])
    out64 = torch.empty_like(wl["pos"])
    f64 = W.advect_ftle(f64field, wl["pos"], positions_out=out64, **kw)
    err = _range_err(f32, f64)
    print(f"[gyre512] FP32 vs FP64 over {wl['n_seeds']} seeds: max|err|/range = {err:.3e}", flush=True)
    assert np.isfinite(f64.cpu().numpy()).all() and err <= 1e-3
    n, nb = wl["n_seeds"], 4
    idx = np.sort(np.random.default_rng(3).choice(n, size=256, replace=False))
    sub = _sample(wl["pos"].cpu().numpy(), n, nb, idx)
    vat = O.time_linear_velocity(_gyre_oracle_fields(wl["host"]["snaps"], 512, 256), 0.0, 0.1)
    exp_ftle, exp_pos = O.run_window_linear(vat, sub, 0.0, wl["h"], wl["steps"], "rk4")
    np.testing.assert_array_equal(_sample(out64.cpu().numpy(), n, nb, idx), exp_pos)
    np.testing.assert_allclose(f64.cpu().numpy()[idx], exp_ftle, rtol=1e-12, atol=1e-14)
    rng = np.ptp(f64.cpu().numpy())
    assert np.max(np.abs(f32.cpu().numpy()[idx] - exp_ftle)) <= 1e-3 * rng


def test_abc128_full_window_fp32_fp64_oracle():
    import bench
    from pyftle_amd import window as W

    wl = bench.build_abc128(torch)
    kw = dict(method="rk4", time_mode="steady", n_steps=wl["steps"], h=wl["h"], map_period=wl["map_period"])
    f32 = W.advect_ftle(wl["field"], wl["pos"], order="morton", **kw)
    f64field = _fp64_twin(wl["field"], [wl["host"]["vel"]])
    out64 = torch.empty_like(wl["pos"])
    f64 = W.advect_ftle(f64field, wl["pos"], positions_out=out64, **kw)
    err = _range_err(f32, f64)
    print(f"[abc128] FP32 vs FP64 over {wl['n_seeds']} seeds, 1000 steps: max|err|/range = {err:.3e}",
          flush=True)
    assert np.isfinite(f64.cpu().numpy()).all() and err <= 1e-3
    # FP64 kernel against the oracle on 64 seeds (periodic trilinear, 1000 RK4 steps)
    n, nb, nn, sp = wl["n_seeds"], 6, wl["host"]["n"], wl["host"]["sp"]
    idx = np.sort(np.random.default_rng(5).choice(n, size=64, replace=False))
    sub = _sample(wl["pos"].cpu().numpy(), n, nb, idx)
    gf = O.GridField.__new__(O.GridField)
    gf.dim, gf.mins, gf.deltas, gf.periodic = 3, [0.0] * 3, [sp] * 3, (True, True, True)
    gf.v = [np.ascontiguousarray(wl["host"]["vel"][c].reshape(nn, nn, nn).transpose(2, 1, 0)) for c in range(3)]
    exp_ftle, exp_pos = O.run_window_linear(lambda t: gf.interpolate, sub, 0.0, wl["h"], wl["steps"], "rk4")
    np.testing.assert_array_equal(_sample(out64.cpu().numpy(), n, nb, idx), exp_pos)
    np.testing.assert_allclose(f64.cpu().numpy()[idx], exp_ftle, rtol=1e-11, atol=1e-13)
    rng = np.ptp(f64.cpu().numpy())
    assert np.max(np.abs(f32.cpu().numpy()[idx] - exp_ftle)) <= 1e-3 * rng


def test_batch64_full_table_one_launch():
    import bench
    from pyftle_amd import window as W
    from pyftle_amd.batch import BatchConfig, BatchPlan

    table = json.load(open(GOLDEN))
    assert table == bench.batch64_table(), "the committed table is the sampled one"
    wl = bench.build_batch64(torch)
    wins = wl["windows"]
    plan = BatchPlan(wins, method="rk4", time_mode="linear",
                     cfg=BatchConfig(groups_per_block=64, blocks_per_cu=4, use_graph=1))
    print(f"[batch64] {plan.info()}, {wl['seed_steps']:.3e} seed-steps", flush=True)
    plan.run()
    torch.cuda.synchronize()
    sizes = [w.positions.shape[0] // 4 for w in wins]
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)),
                    max(range(len(wins)), key=lambda i: wins[i].n_steps), 17, 42})
    for i in picks:
        w, row = wins[i], table[i]
        single = W.advect_ftle(w.field, w.positions, method="rk4", time_mode="linear", n_steps=w.n_steps,
                               h=w.h, t_start=w.t_start, map_period=w.map_period,
                               cfg=W.LaunchConfig(groups_per_block=64))
        assert torch.equal(single, w.ftle), f"field {i}: batch differs from its single launch"
        # FP64 twin of the field (host snapshots of the same flow) against the FP32 batch and the oracle
        nx, ny = row["grid"]
        n_snap = w.field.n_snapshots
        x, y = np.linspace(0.0, 2.0, nx), np.linspace(0.0, 1.0, ny)
        X, Y = np.meshgrid(x, y)
        snaps = []
        pts = np.stack([X.ravel(), Y.ravel()], axis=1)
        for k in range(n_snap):
            snaps.append(np.ascontiguousarray(
                flows.double_gyre(0.1 * k, pts, eps=row["eps"], omega=row["omega"]).T))
        f64field = _fp64_twin(w.field, snaps)
        out64 = torch.empty_like(w.positions)
        f64 = W.advect_ftle(f64field, w.positions, method="rk4", time_mode="linear", n_steps=w.n_steps,
                            h=w.h, t_start=w.t_start, map_period=w.map_period, positions_out=out64)
        err = _range_err(w.ftle, f64)
        print(f"[batch64] field {i} grid {nx}x{ny} seeds {sizes[i]} steps {w.n_steps}: "
              f"FP32 batch vs FP64 max|err|/range = {err:.3e}", flush=True)
        assert err <= 1e-3
        n = sizes[i]
        idx = np.sort(np.random.default_rng(i).choice(n, size=min(n, 32), replace=False))
        sub = _sample(w.positions.cpu().numpy(), n, 4, idx)
        vat = O.time_linear_velocity(_gyre_oracle_fields(snaps, nx, ny), 0.0, 0.1)
        exp_ftle, exp_pos = O.run_window_linear(vat, sub, w.t_start, w.h, w.n_steps, "rk4")
        np.testing.assert_array_equal(_sample(out64.cpu().numpy(), n, 4, idx), exp_pos)
        np.testing.assert_allclose(f64.cpu().numpy()[idx], exp_ftle, rtol=1e-12, atol=1e-14)
        del f64field
    plan.close()
