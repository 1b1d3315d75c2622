"""GPU: the two surrogate configurations at their network, grid and seed shapes, 4 members.

* surrogate2d -- double gyre on 512x256 with 201 snapshots over t in [0, 20]; 64 Fourier
  frequencies on (x, y, t0), 5 x 256 SiLU, 2 outputs; 4M RK4 trajectories (T = 10, h = 0.05,
  t0 ~ U[0, 10)), 200 epochs; seeds 2048 x 1024 at window starts t0 in {0, 2.5, 5, 7.5};
* surrogate3d -- unsteady ABC (A(t) = sqrt3 + 0.5 sin(pi t / 5)) on a periodic 192^3 grid, 64
  snapshots per period; 64 frequencies on (x, y, z, t0), 6 x 512 SiLU, 3 outputs; 4M
  trajectories (T = 5, h = 0.01), 100 epochs; seeds 256^3, judged by RK4 on a 1% held-out
  subset (rng seed 1).

Checked: the ensemble's FTLE (mean Jacobian) within 2% RMSE of the field range against RK4;
a forced fallback (negative tolerance) equal to RK4 bit for bit; the specified gate (FTLE
disagreement 1e-3 of the range) giving a composite field within the bar; the fastest gated
setting on all four members (displacement gate with one Jacobian member: 0.03 in 2D, 1.0 in 3D,
domain units) within the bar, its fallback seeds equal to RK4; the gate on the first two members
(0.03 and 0.1 in 2D, 0.3 and 1.0 in 3D: the bench's fastest settings within the bar) within the
bar at every window start, its fallback seeds equal to RK4.
"""

import math

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pyftle_amd import surrogate

    surrogate._lib()


def test_surrogate2d_config_four_members():
    import bench
    from pyftle_amd import surrogate as S
    from pyftle_amd import surrogate_ftle as SF

    wl = bench.build_surrogate2d(torch)
    field = wl["field"]
    X, D = SF.make_training_data(field, wl["n_traj"], wl["T"], wl["h"], (0.0, 10.0), (0.0, 0.0), (2.0, 1.0),
                                 n_bins=500)
    ens = SF.fit_ensemble(X, D, S.MLPSpec(sigma=1.0), members=4, epochs=wl["epochs"], batch=16384, lr=3e-3,
                          log=lambda m: print(f"[surrogate2d] {m}", flush=True))
    del X, D
    nsx, nsy = wl["nsx"], wl["nsy"]
    xs = (torch.arange(nsx, device="cuda", dtype=torch.float64) + 0.5) * (2.0 / nsx)
    ys = (torch.arange(nsy, device="cuda", dtype=torch.float64) + 0.5) * (1.0 / nsy)
    cy, cx = torch.meshgrid(ys, xs, indexing="ij")
    seeds = torch.stack([cx.reshape(-1), cy.reshape(-1)], 1).contiguous()
    sf = SF.SurrogateFTLE(ens, field, wl["T"], wl["h"], gate="ftle", tol=1e-3)
    sf_d = SF.SurrogateFTLE(ens, field, wl["T"], wl["h"], gate="displacement", tol=0.03, jacobian_members=1)
    # the gate on the ensemble's first two members (the bench's fastest settings within the bar)
    two = S.FlowMapEnsemble(ens.members[:2])
    sf_2 = {tol: SF.SurrogateFTLE(two, field, wl["T"], wl["h"], gate="displacement", tol=tol, jacobian_members=1)
            for tol in (0.03, 0.1)}
    for t0 in (0.0, 2.5, 5.0, 7.5):
        ref = sf.rk4_ftle(seeds, t0)
        f, dis = sf.predict(seeds, t0)
        err = SF.relative_rmse(f, ref)
        res = sf.ftle(seeds, t0)
        err_c = SF.relative_rmse(res.ftle, ref)
        fb = res.fallback.float().mean().item()
        print(f"[surrogate2d] t0 {t0}: ensemble FTLE rmse/range {err:.4f}; gate 1e-3: fallback {fb:.3f}, "
              f"rmse/range {err_c:.4f}", flush=True)
        assert err <= 0.02 and err_c <= err + 1e-12
        assert torch.equal(res.ftle[res.fallback], ref[res.fallback])
        # the fastest gated setting of the bench: displacement gate 0.03, one Jacobian member
        res_d = sf_d.ftle(seeds, t0)
        err_d = SF.relative_rmse(res_d.ftle, ref)
        print(f"[surrogate2d] t0 {t0}: displacement gate 0.03: fallback {res_d.fallback.float().mean().item():.4f}, "
              f"rmse/range {err_d:.4f}", flush=True)
        assert err_d <= 0.02 and torch.equal(res_d.ftle[res_d.fallback], ref[res_d.fallback])
        for tol, s2 in sf_2.items():
            res_2 = s2.ftle(seeds, t0)
            err_2 = SF.relative_rmse(res_2.ftle, ref)
            print(f"[surrogate2d] t0 {t0}: 2-member displacement gate {tol:g}: fallback "
                  f"{res_2.fallback.float().mean().item():.4f}, rmse/range {err_2:.4f}", flush=True)
            assert err_2 <= 0.02 and torch.equal(res_2.ftle[res_2.fallback], ref[res_2.fallback])
    sf.tol = -1.0
    forced = sf.ftle(seeds, 2.5)
    assert bool(forced.fallback.all()) and torch.equal(forced.ftle, sf.rk4_ftle(seeds, 2.5))


def test_surrogate3d_config_four_members():
    import bench
    from pyftle_amd import surrogate as S
    from pyftle_amd import surrogate_ftle as SF

    wl = bench.build_surrogate3d(torch)
    field = wl["field"]
    L = 2.0 * math.pi
    X, D = SF.make_training_data(field, wl["n_traj"], wl["T"], wl["h"], (-1.25, 11.25), (0.0,) * 3, (L,) * 3,
                                 n_bins=400)
    spec = S.MLPSpec(d_in=4, d_out=3, n_freq=64, width=512, depth=6, sigma=0.5, scale=(0.5, 0.5, 0.5, 0.1),
                     out_scale=2.0)
    ens = SF.fit_ensemble(X, D, spec, members=4, epochs=wl["epochs"], batch=16384, lr=2e-3,
                          log=lambda m: print(f"[surrogate3d] {m}", flush=True))
    del X, D
    ns = wl["ns"]
    sax = (torch.arange(ns, device="cuda", dtype=torch.float64) + 0.5) * (L / ns)
    cz, cy, cx = torch.meshgrid(sax, sax, sax, indexing="ij")
    seeds = torch.stack([cx.reshape(-1), cy.reshape(-1), cz.reshape(-1)], 1).contiguous()
    del cx, cy, cz
    g = torch.Generator(device="cuda").manual_seed(1)
    idx = torch.randperm(seeds.shape[0], generator=g, device="cuda")[: seeds.shape[0] // 100].sort().values
    held = seeds[idx].contiguous()
    sf = SF.SurrogateFTLE(ens, field, wl["T"], wl["h"], gate="ftle", tol=1e-3, chunk=1 << 19)
    sf_d = SF.SurrogateFTLE(ens, field, wl["T"], wl["h"], gate="displacement", tol=1.0, jacobian_members=1,
                            chunk=1 << 19)
    # the gate on the ensemble's first two members (the bench's fastest settings within the bar)
    two = S.FlowMapEnsemble(ens.members[:2])
    sf_2 = {tol: SF.SurrogateFTLE(two, field, wl["T"], wl["h"], gate="displacement", tol=tol, jacobian_members=1,
                                  chunk=1 << 19) for tol in (0.3, 1.0)}
    for t0 in wl["t0s"]:
        ref = sf.rk4_ftle(held, t0)
        f, dis = sf.predict(seeds, t0)  # the whole seed grid (the gate's range is the window's)
        f, dis = f[idx], dis[idx]
        err = SF.relative_rmse(f, ref)
        bad = dis > sf.tol
        comp = torch.where(bad, ref, f)
        err_c = SF.relative_rmse(comp, ref)
        print(f"[surrogate3d] t0 {t0}: ensemble FTLE rmse/range {err:.4f} on the held-out 1%; gate 1e-3: "
              f"fallback {bad.float().mean().item():.3f}, rmse/range {err_c:.4f}", flush=True)
        assert err <= 0.02 and err_c <= err + 1e-12
        # the fastest gated setting of the bench: displacement gate 1.0 (domain units), one
        # Jacobian member, on the held-out seeds (probe plan)
        res_d = sf_d.ftle(held, t0)
        err_d = SF.relative_rmse(res_d.ftle, ref)
        print(f"[surrogate3d] t0 {t0}: displacement gate 1.0: fallback {res_d.fallback.float().mean().item():.4f}, "
              f"rmse/range {err_d:.4f}", flush=True)
        assert err_d <= 0.02 and torch.equal(res_d.ftle[res_d.fallback], ref[res_d.fallback])
        for tol, s2 in sf_2.items():
            res_2 = s2.ftle(held, t0)
            err_2 = SF.relative_rmse(res_2.ftle, ref)
            print(f"[surrogate3d] t0 {t0}: 2-member displacement gate {tol:g}: fallback "
                  f"{res_2.fallback.float().mean().item():.4f}, rmse/range {err_2:.4f}", flush=True)
            assert err_2 <= 0.02 and torch.equal(res_2.ftle[res_2.fallback], ref[res_2.fallback])
    sf.tol = -1.0
    forced = sf.ftle(held, wl["t0s"][1])
    assert bool(forced.fallback.all()) and torch.equal(forced.ftle, sf.rk4_ftle(held, wl["t0s"][1]))
