"""GPU: the per-cell record layouts (CELL, CELL_PAIR) against the node layouts (SINGLE, PAIR).

A cell record holds the same FP32 values as the 2^dim node records of the cell and the
kernel does the same arithmetic on them, so final positions and FTLE are bit-identical:
periodic and bounded axes, particles that leave the domain (zero velocity outside), every
integrator, steady / frozen / time-linear snapshots, single launches and batches.  The
layouts hold FP32 GridInterp fields, read by the FP32 advection kernels and the gather call;
other uses are refused.
"""

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


def _field(vels, pts, shape, layout, periodic=None, dt_snap=0.1, precision="float32"):
    from pyftle_amd.field import DeviceField, GridGeometry

    kw = {} if periodic is None else {"periodic": periodic}
    f = DeviceField(GridGeometry.from_reference(shape, pts, **kw), dtype=precision, layout=layout,
                    ring_slots=len(vels), n_snapshots=len(vels), t0=0.0, dt_snap=dt_snap)
    for k, v in enumerate(vels):
        f.pack(k, v)
    return f


def _run(f, x0, **kw):
    from pyftle_amd import window as W

    pos = torch.as_tensor(x0, device="cuda")
    out = torch.empty_like(pos)
    ftle = W.advect_ftle(f, pos, positions_out=out, **kw)
    return ftle, out


def _seeds_2d_with_outside():
    """Seed groups over the domain [0, 2] x [0, 1] and a band around it: some start
    outside, some on the upper boundary, some leave during the window."""
    x0 = H.seeds_2d(23, 11, lo=(-0.1, -0.05), hi=(2.1, 1.0))
    return x0


@pytest.mark.parametrize("method", ["euler", "ab2", "rk4"])
def test_cell_layout_2d_bounded_bit_identical(method):
    from pyftle_amd import _native as N

    vels, pts = H.double_gyre_window(nx=48, ny=24, n_snap=7)
    x0 = _seeds_2d_with_outside()
    node = _field(vels, pts, (48, 24), N.LAYOUT_SINGLE)
    cell = _field(vels, pts, (48, 24), N.LAYOUT_CELL)
    for kw in (dict(time_mode="frozen", n_steps=6, h=0.1, map_period=0.5),
               dict(time_mode="steady", n_steps=40, h=-0.05, map_period=2.0)):
        f0, p0 = _run(node, x0, method=method, **kw)
        f1, p1 = _run(cell, x0, method=method, **kw)
        assert torch.equal(p1, p0)
        assert torch.equal(f1, f0)
        assert not torch.equal(p0, torch.as_tensor(x0, device="cuda"))  # the particles moved


def test_cell_pair_layout_2d_time_linear_bit_identical():
    from pyftle_amd import _native as N

    vels, pts = H.double_gyre_window(nx=64, ny=32, n_snap=11)
    x0 = _seeds_2d_with_outside()
    node = _field(vels, pts, (64, 32), N.LAYOUT_PAIR)
    cell = _field(vels, pts, (64, 32), N.LAYOUT_CELL_PAIR)
    for t_start, h in ((0.0, 0.05), (0.93, -0.04)):
        kw = dict(method="rk4", time_mode="linear", n_steps=20, h=h, t_start=t_start, map_period=20 * abs(h))
        f0, p0 = _run(node, x0, **kw)
        f1, p1 = _run(cell, x0, **kw)
        assert torch.equal(p1, p0)
        assert torch.equal(f1, f0)


@pytest.mark.parametrize("mode", ["steady", "linear", "frozen"])
def test_cell_layouts_3d_periodic_bit_identical(mode):
    """The FP32 kernel keeps each lane's last cell record in registers, keyed by record and
    time slot: steady (one slot), time-linear pairs (slots change with the stage times) and
    frozen snapshots (a new slot every step) all match the node layouts bit for bit."""
    from pyftle_amd import _native as N

    n = 20
    vels, pts = H.abc_window(n=n, n_snap=1 if mode == "steady" else 5, snap_dt=0.25, unsteady=mode != "steady")
    per = (True, True, True)
    lay0, lay1 = (N.LAYOUT_PAIR, N.LAYOUT_CELL_PAIR) if mode == "linear" else (N.LAYOUT_SINGLE, N.LAYOUT_CELL)
    node = _field(vels, pts, (n, n, n), lay0, periodic=per, dt_snap=0.25)
    cell = _field(vels, pts, (n, n, n), lay1, periodic=per, dt_snap=0.25)
    x0 = H.seeds_3d(9, lo=0.05, hi=6.2)  # groups across the periodic seams
    if mode == "frozen":  # one snapshot per step, five steps
        kw = dict(method="rk4", time_mode="frozen", n_steps=5, h=0.2, map_period=1.0)
    else:
        kw = dict(method="rk4", time_mode=mode, n_steps=80, h=0.0125, t_start=0.05, map_period=1.0)
    f0, p0 = _run(node, x0, **kw)
    f1, p1 = _run(cell, x0, **kw)
    assert torch.equal(p1, p0)
    assert torch.equal(f1, f0)
    assert np.isfinite(f0.cpu().numpy()).all()


@pytest.mark.parametrize("method", ["euler", "ab2", "rk4"])
def test_cell_layout_plain_advection_bit_identical(method):
    """pyftle_advect (the per-step integrator entry point, no FTLE) on 2D periodic and 3D
    bounded fields, FP64 positions."""
    from pyftle_amd import _native as N
    from pyftle_amd import window as W

    cases = []
    vels, pts = H.double_gyre_window(nx=40, ny=20, n_snap=3)
    cases.append((vels, pts, (40, 20), (True, False), H.seeds_2d(17, 9)))
    vels3, pts3 = H.abc_window(n=12, endpoint=True)
    cases.append((vels3, pts3, (12, 12, 12), None, H.seeds_3d(5, lo=0.3, hi=6.0)))
    for vels_, pts_, shape, per, x0 in cases:
        lay0, lay1 = N.LAYOUT_SINGLE, N.LAYOUT_CELL
        out = []
        for lay in (lay0, lay1):
            f = _field(vels_, pts_, shape, lay, periodic=per)
            pos = torch.as_tensor(x0, device="cuda").clone()
            vprev = torch.zeros_like(pos) if method == "ab2" else None
            for k in range(2):  # two calls: AB2 carries its previous velocity across them
                W.advect(f, pos, method=method, time_mode="steady", n_steps=7, h=0.03, first=k % len(vels_),
                         vprev=vprev, have_prev=(method == "ab2" and k > 0))
            out.append(pos)
        assert torch.equal(out[1], out[0])


def test_cell_pair_batch_matches_node_layout():
    from pyftle_amd import _native as N
    from pyftle_amd.batch import BatchPlan, Window

    specs = [((32, 16), 9, (7, 3), 12, 0.05, 0.0), ((48, 24), 13, (21, 9), 25, -0.05, 0.6)]
    outs = {}
    for lay in (N.LAYOUT_PAIR, N.LAYOUT_CELL_PAIR):
        wins = []
        for grid, n_snap, (sx, sy), steps, h, t0 in specs:
            vels, pts = H.double_gyre_window(nx=grid[0], ny=grid[1], n_snap=n_snap)
            f = _field(vels, pts, grid, lay)
            pos = torch.as_tensor(H.seeds_2d(sx, sy), device="cuda")
            wins.append(Window(field=f, positions=pos, n_steps=steps, h=h, t_start=t0, map_period=steps * abs(h)))
        plan = BatchPlan(wins, method="rk4", time_mode="linear")
        plan.run()
        torch.cuda.synchronize()
        outs[lay] = [w.ftle.clone() for w in wins]
        plan.close()
    for a, b in zip(outs[N.LAYOUT_PAIR], outs[N.LAYOUT_CELL_PAIR]):
        assert torch.equal(a, b)


def test_cell_layouts_refuse_other_uses():
    from pyftle_amd import _native as N
    from pyftle_amd import window as W

    vels, pts = H.double_gyre_window(nx=16, ny=8, n_snap=2)
    with pytest.raises(RuntimeError, match="CELL layouts hold FP32"):
        _field(vels, pts, (16, 8), N.LAYOUT_CELL, precision="float64")
    f = _field(vels, pts, (16, 8), N.LAYOUT_CELL)
    x0 = torch.as_tensor(H.seeds_2d(3, 2), device="cuda")
    with pytest.raises(RuntimeError, match="compute must be NATIVE"):
        W.advect_ftle(f, x0, method="rk4", time_mode="steady", n_steps=2, h=0.1, map_period=0.2,
                      cfg=W.LaunchConfig(compute=N.COMPUTE_MIXED))
    with pytest.raises(RuntimeError, match="LINEAR time needs"):
        W.advect_ftle(f, x0, method="rk4", time_mode="linear", n_steps=2, h=0.1, map_period=0.2)


@pytest.mark.parametrize("pts_dtype", [torch.float64, torch.float32])
def test_cell_layouts_field_sample_bit_identical(pts_dtype):
    """The gather call reads the cell records too, with the node layouts' result bit for bit:
    2D time-linear pairs on a bounded grid (points inside, on the edges and outside) and a
    3D periodic field (points beyond the seams wrap)."""
    from pyftle_amd import _native as N

    vels, pts = H.double_gyre_window(nx=33, ny=17, n_snap=4)
    node = _field(vels, pts, (33, 17), N.LAYOUT_PAIR)
    cell = _field(vels, pts, (33, 17), N.LAYOUT_CELL_PAIR)
    rng = np.random.default_rng(3)
    q = np.concatenate([rng.uniform((-0.2, -0.1), (2.2, 1.1), size=(500, 2)),
                        [[0.0, 0.0], [2.0, 1.0], [2.0, 0.5], [1.0, 1.0]]])
    q = torch.as_tensor(q, device="cuda", dtype=pts_dtype)
    for t in (0.0, 0.13, 0.29):
        a = node.sample(q, time_mode=N.TIME_LINEAR, t=t)
        b = cell.sample(q, time_mode=N.TIME_LINEAR, t=t)
        assert torch.equal(a, b)
    assert (a != 0).any() and (a == 0).any()  # inside and outside points

    n = 12
    vels, pts = H.abc_window(n=n, n_snap=1)
    per = (True, True, True)
    node = _field(vels, pts, (n, n, n), N.LAYOUT_SINGLE, periodic=per)
    cell = _field(vels, pts, (n, n, n), N.LAYOUT_CELL, periodic=per)
    q = torch.as_tensor(rng.uniform(-1.0, 7.5, size=(700, 3)), device="cuda", dtype=pts_dtype)
    assert torch.equal(node.sample(q, time_mode=N.TIME_STEADY), cell.sample(q, time_mode=N.TIME_STEADY))


@pytest.mark.parametrize("case", ["3d-steady-periodic", "2d-linear-pair"])
def test_counting_launch_identical_and_bounded(case):
    """The counting copy of the kernel (pyftle_advect_ftle_count_loads) gives the same
    positions and FTLE bit for bit, and its count of wave-level record loads lies between
    one record per wave (every lane's first sample loads) and every sample of every wave
    loading (4 RK4 stages per step, 6 (3D) or 4 (2D pair) 16-byte loads per record)."""
    from pyftle_amd import _native as N
    from pyftle_amd import window as W

    if case.startswith("3d"):
        vels, pts = H.abc_window(n=20)
        f = _field(vels, pts, (20, 20, 20), N.LAYOUT_CELL, periodic=(True, True, True))
        x0 = H.seeds_3d(9)
        kw = dict(method="rk4", time_mode="steady", n_steps=50, h=0.05, map_period=2.5)
        nb, per_record = 6, 6
    else:
        vels, pts = H.double_gyre_window(nx=64, ny=32, n_snap=11)
        f = _field(vels, pts, (64, 32), N.LAYOUT_CELL_PAIR)
        x0 = _seeds_2d_with_outside()
        kw = dict(method="rk4", time_mode="linear", n_steps=18, h=0.05, t_start=0.05, map_period=0.9)
        nb, per_record = 4, 4
    cfg = W.LaunchConfig(groups_per_block=32)
    f0, p0 = _run(f, x0, cfg=cfg, **kw)
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    pos = torch.as_tensor(x0, device="cuda")
    out = torch.empty_like(pos)
    f1 = W.advect_ftle(f, pos, positions_out=out, cfg=cfg, count_loads=cnt, **kw)
    assert torch.equal(f1, f0) and torch.equal(out, p0)
    n_groups = x0.shape[0] // nb
    waves = -(-n_groups // 32) * (-(-(32 * nb) // 64))
    loads = int(cnt.item())
    print(f"[count {case}] {loads} wave loads, {waves} waves, {kw['n_steps']} steps", flush=True)
    assert waves * per_record <= loads <= waves * kw["n_steps"] * 4 * per_record
    with pytest.raises(RuntimeError):  # node records are not counted
        node = _field(vels, pts, (20, 20, 20) if case.startswith("3d") else (64, 32),
                      N.LAYOUT_SINGLE if case.startswith("3d") else N.LAYOUT_PAIR,
                      periodic=(True, True, True) if case.startswith("3d") else None)
        W.advect_ftle(node, pos, cfg=cfg, count_loads=cnt, **kw)


def test_probe_load_rate_is_positive():
    from pyftle_amd import window as W

    r = W.probe_load_rate(blocks=512, iters=1024, repeats=2)
    print(f"[probe] {r['wave_loads_per_s']:.3e} wave loads/s = {r['bytes_per_s'] / 1e12:.1f} TB/s", flush=True)
    assert r["wave_loads_per_s"] > 0


def test_probe_issue_rates_cover_every_mode():
    """The issue-rate probe runs all eight modes; an MFMA 16x16x32 bf16 cannot issue faster
    than the 16 cycles its 16384 flops take at 1024 flops per cycle per SIMD."""
    from pyftle_amd import window as W

    r = W.probe_issue_rates(iters=500)
    print(f"[probe] {r}", flush=True)
    assert set(r) == set(W.ISSUE_MODES.values())
    for name, row in r.items():
        if "cycles_per_mfma" in row:
            assert row["cycles_per_mfma"] >= 15.0, name
        if "cycles_per_valu" in row:
            assert row["cycles_per_valu"] > 0.0, name
