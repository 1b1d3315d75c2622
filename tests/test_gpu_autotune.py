"""GPU: the launch-shape autotuners.

``tuning.autotune`` times every single-window candidate (seeds per work-group, lane order,
XCD remap) and stores the fastest per shape class; ``batch.autotune_batch`` does the same for
batched launches.  The tuned shapes change only the speed: every candidate gives the same
FTLE bit for bit.  The cache is redirected to a temporary file.
"""

import json

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


@pytest.fixture
def tmp_cache(tmp_path, monkeypatch):
    from pyftle_amd import tuning

    path = tmp_path / "tuning_cache.json"
    monkeypatch.setattr(tuning, "CACHE", str(path))
    tuning._memo["mtime"] = None  # forget the committed cache's entries
    yield path
    tuning._memo["mtime"] = None  # re-read the committed cache on next use


def _gyre_pair_field(nx=64, ny=32, n_snap=11):
    from pyftle_amd import _native as N
    from pyftle_amd.field import DeviceField, GridGeometry

    vels, pts = H.double_gyre_window(nx=nx, ny=ny, n_snap=n_snap)
    f = DeviceField(GridGeometry.from_reference((nx, ny), pts), dtype="float32", layout=N.LAYOUT_PAIR,
                    ring_slots=n_snap, n_snapshots=n_snap, t0=0.0, dt_snap=0.1)
    for k, v in enumerate(vels):
        f.pack(k, v)
    return f


def test_autotune_single_window_picks_a_candidate_and_saves_it(tmp_cache):
    from pyftle_amd import tuning
    from pyftle_amd import window as W

    f = _gyre_pair_field()
    x0 = torch.as_tensor(H.seeds_2d(40, 20), device="cuda")
    kw = dict(method="rk4", time_mode="linear", n_steps=16, h=0.05, map_period=0.8)
    ref = W.advect_ftle(f, x0, cfg=W.LaunchConfig(), **kw)

    def run(cfg):
        return W.advect_ftle(f, x0, cfg=cfg, **kw)

    key = tuning.shape_class(2, x0.shape[0] // 4, "linear")
    best, table = tuning.autotune(run, 2, key, repeats=2)
    cands = tuning.candidates(2)
    assert len(table) == len(cands)
    assert [ms for ms, _ in table] == sorted(ms for ms, _ in table)
    assert vars(best) == table[0][1]
    for cfg in cands:  # every shape gives the same field
        assert torch.equal(run(cfg), ref)
    saved = json.loads(tmp_cache.read_text())[key]
    assert saved["groups_per_block"] == best.groups_per_block and saved["xcd_remap"] == best.xcd_remap
    got = tuning.best_config(key)
    assert (got.groups_per_block, got.lane_order, got.xcd_remap) == (best.groups_per_block, best.lane_order,
                                                                      best.xcd_remap)
    # a call without a launch config now takes the tuned shape, with the same result
    assert W.tuned_config(f, x0.shape[0] // 4, "linear").groups_per_block == best.groups_per_block
    assert torch.equal(W.advect_ftle(f, x0, **kw), ref)


def test_autotune_batch_returns_the_fastest_configuration(tmp_cache):
    from pyftle_amd import tuning
    from pyftle_amd.batch import BatchPlan, Window, autotune_batch

    wins = []
    for nx, ny, sx, sy, steps in ((32, 16, 9, 5, 10), (64, 32, 30, 14, 16)):
        f = _gyre_pair_field(nx, ny)
        pos = torch.as_tensor(H.seeds_2d(sx, sy), device="cuda")
        wins.append(Window(field=f, positions=pos, n_steps=steps, h=0.05, map_period=steps * 0.05))
    best, table = autotune_batch(wins, repeats=1)
    assert [ms for ms, _ in table] == sorted(ms for ms, _ in table)
    assert vars(best) == table[0][1]
    plan = BatchPlan(wins, cfg=best)
    plan.run()
    torch.cuda.synchronize()
    tuned = [w.ftle.clone() for w in wins]
    plan.close()
    plan = BatchPlan(wins)  # default configuration
    plan.run()
    torch.cuda.synchronize()
    for w, t in zip(wins, tuned):
        assert torch.equal(w.ftle, t)
    plan.close()
    tuning.save_batch_config("small-batch", best, table[0][0])
    assert vars(tuning.best_batch_config("small-batch")) == vars(best)
