"""GPU: the learned flow-map surrogate against plain RK4 (its fallback and accuracy reference).

* trained on RK4 trajectories of a small double gyre, the ensemble's FTLE is
  within 10% (RMSE / range) of RK4;
* on a longer window one trained member gains over the method without it
  (plain RK4 on the same seeds) in time per field within the 2% RMSE bar (measured
  2.8x at 1.4-1.7%); the full-size measurement is ``bench.py --workload surrogate2d``;
* the probe plan keeps the per-seed gate: seeds from the ensemble passed it, the rest
  are RK4 bit for bit;
* a member perturbed on purpose disagrees everywhere, so every seed falls back
  and the field equals the RK4 field exactly; the FTLE gate's disagreement is zero
  for identical members and flags a member whose Jacobian differs;
* on a linear flow the JVP Jacobian of a network fitted to the exact
  displacement gives the closed-form FTLE.
"""

import math

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyftle_amd import surrogate

    surrogate._lib()


@pytest.fixture(scope="module")
def gyre_setup():
    from pyftle_amd import _native as N
    from pyftle_amd import surrogate as S
    from pyftle_amd import surrogate_ftle as SF
    from pyftle_amd.field import DeviceField, GridGeometry

    vels, pts = H.double_gyre_window(nx=128, ny=64, n_snap=41, snap_dt=0.1)
    f = DeviceField(GridGeometry.from_reference((128, 64), pts), dtype="float32", layout=N.LAYOUT_PAIR,
                    ring_slots=41, n_snapshots=41, t0=0.0, dt_snap=0.1)
    for k, v in enumerate(vels):
        f.pack(k, v)
    T, h = 2.0, 0.05
    X, D = SF.make_training_data(f, 262144, T, h, (0.0, 2.0), (0.05, 0.05), (1.95, 0.95), n_bins=64)
    spec = S.MLPSpec(sigma=1.0, width=128, depth=3, scale=(0.5, 1.0, 0.5))
    ens = SF.fit_ensemble(X, D, spec, members=2, epochs=25, batch=8192, lr=3e-3)
    return f, ens, T, h


def _seeds(nsx=128, nsy=64):
    xs = (torch.arange(nsx, device="cuda", dtype=torch.float64) + 0.5) * (1.8 / nsx) + 0.1
    ys = (torch.arange(nsy, device="cuda", dtype=torch.float64) + 0.5) * (0.8 / nsy) + 0.1
    cy, cx = torch.meshgrid(ys, xs, indexing="ij")
    return torch.stack([cx.reshape(-1), cy.reshape(-1)], 1).contiguous()


def test_surrogate_ftle_close_to_rk4(gyre_setup):
    from pyftle_amd import surrogate_ftle as SF

    f, ens, T, h = gyre_setup
    sf = SF.SurrogateFTLE(ens, f, T, h, tol=1e9)  # no fallback: the network alone
    seeds = _seeds()
    for t0 in (0.3, 1.4):
        res = sf.ftle(seeds, t0)
        ref = sf.rk4_ftle(seeds, t0)
        err = SF.relative_rmse(res.ftle, ref)
        print(f"[gpu] t0={t0} surrogate FTLE rmse/range {err:.4f}", flush=True)
        assert err < 0.10
        assert not bool(res.fallback.any())


def test_fallback_recomputes_disagreeing_seeds_with_rk4(gyre_setup):
    from pyftle_amd import surrogate_ftle as SF
    from pyftle_amd.surrogate import FlowMapEnsemble

    f, ens, T, h = gyre_setup
    bad = ens.members[1]
    saved = bad.params.clone()
    o, shp, ob = bad.out_off
    bad.params[ob:ob + shp[0]] += 0.5  # shift one member's output: it disagrees everywhere
    bad.sync_bf16()
    try:
        # a constant output shift leaves the Jacobians alone: the displacement gate sees it
        sf = SF.SurrogateFTLE(FlowMapEnsemble(ens.members), f, T, h, tol=1e-3, gate="displacement")
        seeds = _seeds(64, 32)
        res = sf.ftle(seeds, 0.7)
        assert bool(res.fallback.all())
        assert torch.equal(res.ftle, sf.rk4_ftle(seeds, 0.7))
        # several window starts, fallback seeds batched into one launch: the same fields
        batched = sf.ftle_windows(seeds, [0.7, 1.1, 0.2])
        for t0, r in zip([0.7, 1.1, 0.2], batched):
            assert bool(r.fallback.all())
            assert torch.equal(r.ftle, sf.ftle(seeds, t0).ftle)
        sf.tol = 1e9  # no fallback at all
        for t0, r in zip([0.4, 1.3], sf.ftle_windows(seeds, [0.4, 1.3])):
            assert not bool(r.fallback.any())
            assert torch.equal(r.ftle, sf.ftle(seeds, t0).ftle)
    finally:
        bad.params.copy_(saved)
        bad.sync_bf16()


def test_jvp_ftle_on_linear_flow():
    """Saddle u = l x, v = -l y: displacement d(x) = (e^{lT}-1) x, (e^{-lT}-1) y exactly."""
    from pyftle_amd import surrogate as S

    lam, T = 0.4, 1.0
    spec = S.MLPSpec(d_in=3, sigma=0.5, width=128, depth=2, scale=(1.0, 1.0, 1.0))
    m = S.FlowMapMLP(spec, seed=5)
    g = torch.Generator(device="cuda").manual_seed(0)
    X = torch.rand(65536, 3, device="cuda", generator=g) - 0.5
    X[:, 2] = 0.0
    D = torch.stack([(math.exp(lam * T) - 1) * X[:, 0], (math.exp(-lam * T) - 1) * X[:, 1]], 1)
    S.train(m, X, D, epochs=40, batch=8192, lr=3e-3)
    q = torch.rand(4096, 3, device="cuda", generator=g) * 0.6 - 0.3
    q[:, 2] = 0.0
    _, Jd, _ = S.FlowMapEnsemble([m]).predict(q, 2)
    ftle = S.ftle_from_displacement_jacobian(Jd, T).cpu().numpy()
    print(f"[gpu] linear-flow FTLE mean {ftle.mean():.4f} (exact {lam})", flush=True)
    assert abs(np.median(ftle) - lam) < 0.02


def test_surrogate_3d_periodic_pipeline():
    """Unsteady periodic ABC (3 tangents, 3x3 Jacobians, layer-by-layer kernels at width
    128): the fitted ensemble's FTLE tracks RK4, and forced fallback returns RK4 exactly."""
    import math as m_

    from pyftle_amd import _native as N
    from pyftle_amd import surrogate as S
    from pyftle_amd import surrogate_ftle as SF
    from pyftle_amd.field import DeviceField, GridGeometry

    n, L, dts = 32, 2 * m_.pi, 0.25
    geom = GridGeometry.uniform((n, n, n), (0.0,) * 3, (L / n,) * 3, periodic=(True,) * 3)
    f = DeviceField(geom, dtype="float32", layout=N.LAYOUT_PAIR, ring_slots=13, n_snapshots=13, t0=0.0,
                    dt_snap=dts)
    ax = torch.arange(n, dtype=torch.float64, device="cuda") * (L / n)
    Z, Y, X = torch.meshgrid(ax, ax, ax, indexing="ij")
    for k in range(13):
        a = m_.sqrt(3.0) + 0.5 * m_.sin(m_.pi * k * dts / 5.0)
        f.pack(k, torch.stack([a * Z.sin() + Y.cos(), m_.sqrt(2.0) * X.sin() + a * Z.cos(),
                               Y.sin() + m_.sqrt(2.0) * X.cos()]).reshape(3, -1))
    T, h = 1.0, 0.02
    X_, D_ = SF.make_training_data(f, 131072, T, h, (0.0, 2.0), (0.0,) * 3, (L,) * 3, n_bins=16)
    spec = S.MLPSpec(d_in=4, d_out=3, n_freq=64, width=128, depth=3, sigma=0.5, scale=(0.5, 0.5, 0.5, 0.5))
    ens = SF.fit_ensemble(X_, D_, spec, members=2, epochs=20, batch=8192, lr=3e-3)
    g = torch.Generator(device="cuda").manual_seed(1)
    seeds = torch.rand(4096, 3, dtype=torch.float64, device="cuda", generator=g) * L
    sf = SF.SurrogateFTLE(ens, f, T, h, tol=1e9)
    res = sf.ftle(seeds, 1.0)
    ref = sf.rk4_ftle(seeds, 1.0)
    assert bool(torch.isfinite(res.ftle).all())
    err = SF.relative_rmse(res.ftle, ref)
    print(f"[gpu] 3D surrogate FTLE rmse/range {err:.4f}", flush=True)
    assert err < 0.15
    sf.tol = -1.0  # every seed disagrees: the whole field comes from RK4
    res = sf.ftle(seeds, 1.0)
    assert bool(res.fallback.all()) and torch.equal(res.ftle, ref)


def test_trained_network_gains_over_rk4_at_fixed_accuracy():
    """What the trained network buys: on a longer window (T = 5, 100 RK4 steps) one
    fused-kernel member (width 256, 3 hidden layers, 2 JVP tangents) gives the FTLE of
    524,288 seeds faster than the plain RK4 kernel on the same seeds, within 5% RMSE
    of the range.  The network's cost does not grow with the window; RK4's does."""
    from pyftle_amd import _native as N
    from pyftle_amd import surrogate as S
    from pyftle_amd import surrogate_ftle as SF
    from pyftle_amd.field import DeviceField, GridGeometry

    vels, pts = H.double_gyre_window(nx=256, ny=128, n_snap=81, snap_dt=0.1)
    f = DeviceField(GridGeometry.from_reference((256, 128), pts), dtype="float32", layout=N.LAYOUT_PAIR,
                    ring_slots=81, n_snapshots=81, t0=0.0, dt_snap=0.1)
    for k, v in enumerate(vels):
        f.pack(k, v)
    T, h = 5.0, 0.05
    X, D = SF.make_training_data(f, 1 << 20, T, h, (0.0, 3.0), (0.05, 0.05), (1.95, 0.95), n_bins=128)
    spec = S.MLPSpec(sigma=1.0, width=256, depth=3, scale=(0.5, 1.0, 0.3))
    ens = SF.fit_ensemble(X, D, spec, members=1, epochs=30, batch=16384, lr=3e-3)
    assert ens.members[0].fused_supported(2)
    sf = SF.SurrogateFTLE(ens, f, T, h, tol=1e9)
    seeds = _seeds(1024, 512)

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            out = fn()
            b.record()
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        return out, sorted(ms)[reps // 2]

    errs, gains = [], []
    for t0 in (0.5, 2.0):
        sur, t_sur = timed(lambda: sf.predict(seeds, t0)[0])
        ref, t_rk4 = timed(lambda: sf.rk4_ftle(seeds, t0))
        errs.append(SF.relative_rmse(sur, ref))
        gains.append(t_rk4 / t_sur)
        print(f"[gpu] t0={t0}: surrogate {t_sur:.2f} ms, RK4 {t_rk4:.2f} ms, gain {gains[-1]:.2f}x, "
              f"FTLE rmse/range {errs[-1]:.4f}", flush=True)
    assert max(errs) <= 0.02
    assert min(gains) > 1.5


def test_ftle_gate_units_and_forced_fallback(gyre_setup):
    """The FTLE gate: identical members agree exactly (disagreement 0); a member whose
    first hidden layer is scaled disagrees, in units of the window's FTLE range; a negative
    tolerance forces every seed to RK4, bit for bit."""
    from pyftle_amd import surrogate_ftle as SF
    from pyftle_amd.surrogate import FlowMapEnsemble

    f, ens, T, h = gyre_setup
    seeds = _seeds(64, 32)
    same = SF.SurrogateFTLE(FlowMapEnsemble([ens.members[0], ens.members[0]]), f, T, h, gate="ftle")
    fs, dis = same.predict(seeds, 0.5)
    assert float(dis.max()) == 0.0
    sf = SF.SurrogateFTLE(ens, f, T, h, gate="ftle", tol=-1.0)
    res = sf.ftle(seeds, 0.5)
    assert bool(res.fallback.all()) and torch.equal(res.ftle, sf.rk4_ftle(seeds, 0.5))
    bad = ens.members[1]
    saved = bad.params.clone()
    o, shp, ob = bad.offsets[0]
    bad.params[o:o + shp[0] * shp[1]] *= 1.5
    bad.sync_bf16()
    try:
        _, dis2 = SF.SurrogateFTLE(FlowMapEnsemble(ens.members), f, T, h, gate="ftle").predict(seeds, 0.5)
        print(f"[gpu] perturbed member: FTLE disagreement / range median {dis2.median().item():.3e}", flush=True)
        assert float(dis2.median()) > 1e-2
    finally:
        bad.params.copy_(saved)
        bad.sync_bf16()


def test_probe_plan_keeps_the_per_seed_gate(gyre_setup):
    """The probe plan (65,536 seeds, probe every 64th here, 5x5 tiles): every seed returned from
    the ensemble passed its own gate and carries the ensemble's value for that seed; every
    other seed is RK4 bit for bit; tiles sent to RK4 whole are those whose probed seeds
    failed beyond the share; a tolerance nothing passes sends the whole field to RK4 and one
    everything passes leaves the ensemble's field."""
    from pyftle_amd import surrogate_ftle as SF

    f, ens, T, h = gyre_setup
    seeds = _seeds(256, 256)
    plain = SF.SurrogateFTLE(ens, f, T, h, gate="ftle", probe=0)
    f_all, dis_all = plain.predict(seeds, 0.5)
    ref = plain.rk4_ftle(seeds, 0.5)
    tol = float(dis_all.median())
    sf = SF.SurrogateFTLE(ens, f, T, h, gate="ftle", tol=tol, tile_fallback_max=0.5, probe=64)
    assert sf._planned(seeds)
    plan = sf._tiles_of(seeds)
    tid, pidx, n_tiles = plan.tile, plan.probe, plan.n_tiles
    assert n_tiles == 25 and pidx.numel() == seeds.shape[0] // 64
    res = sf.ftle(seeds, 0.5)
    skipped = torch.isnan(res.spread)
    kept = ~res.fallback
    print(f"[gpu] probe plan at the median disagreement: fallback {res.fallback.float().mean().item():.3f}, "
          f"tiles to RK4 {skipped.float().mean().item():.3f}", flush=True)
    assert 0.0 < skipped.float().mean().item() < 1.0
    assert bool((res.spread[kept] <= tol).all())
    assert torch.equal(res.ftle[kept], f_all[kept])
    assert torch.equal(res.ftle[res.fallback], ref[res.fallback])
    assert bool(res.fallback[skipped].all())
    # a skipped tile is one whose probed seeds failed beyond the share
    f_p, dev_p = sf._evaluate(seeds[pidx].contiguous(), 0.5)
    bad_p = sf._normalise(dev_p, sf._range(f_p)) > tol
    frac = torch.bincount(tid[pidx][bad_p], minlength=n_tiles).double() / torch.bincount(
        tid[pidx], minlength=n_tiles).double()
    assert torch.equal((frac > 0.5)[tid], skipped)
    tile_rk4, tile_rk4_host, _ = sf._probe(seeds, [0.5], plan)[0]
    assert torch.equal(tile_rk4.cpu(), tile_rk4_host) and torch.equal(tile_rk4_host, (frac > 0.5).cpu())
    # the batched form equals the single-window form
    for t0, r in zip([0.5, 1.1], sf.ftle_windows(seeds, [0.5, 1.1])):
        one = sf.ftle(seeds, t0)
        assert torch.equal(r.fallback, one.fallback) and torch.equal(r.ftle, one.ftle)
    sf.tol = -1.0
    res = sf.ftle(seeds, 0.5)
    assert bool(torch.isnan(res.spread).all()) and torch.equal(res.ftle, ref)
    sf.tol = 1e9
    res = sf.ftle(seeds, 0.5)
    assert not bool(res.fallback.any()) and torch.equal(res.ftle, f_all)
