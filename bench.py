#!/usr/bin/env python
"""FTLE throughput benchmark on one MI355X.

A step is one FTLE field: the fused advection kernel takes every seed group
through the whole window and its FP64 epilogue writes the FTLE per seed.

Workloads (synthetic, generated deterministically on the fly):
* ``abc128`` (default): steady ABC flow (A=sqrt3, B=sqrt2, C=1) on a periodic
  128^3 FP32 grid over [0, 2pi)^3; seeds on a 128^3 cell-centred grid, 6 axis
  neighbours at delta=1e-3; RK4 h=0.01, 1000 steps, trilinear.
* ``gyre512``: double gyre (A=0.1, eps=0.25, omega=2pi/10) on 512x256 over
  [0,2]x[0,1], 101 snapshots t=0..10 (FP64 on host -> FP32 in HBM, PAIR
  layout); 400x200 cell-centred seeds, 4 neighbours at 1e-3; RK4 h=0.05,
  200 steps, linear in time.

Prints one JSON line (metric, value = FTLE seeds/s, roofline of the advection
kernel against 8 TB/s HBM, accuracy against FP64, CPU baseline).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FTLE seeds/s at fixed accuracy: fused RK4 vs MLP surrogate on 1 MI355X; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

# record layout of the FP32 fields: "node" (one record per grid node, the layout of the
# reference's arrays), "cell" (one record per cell holding its 2^dim corners) or "auto"
# (cell records: a sample is one record, 6 contiguous 16-byte loads in 3D instead of 8
# scattered node records, and the kernels keep each lane's last record in registers, so
# a sample in the same cell as the lane's previous one loads nothing; abc128 226 -> 118 ms,
# gyre512 1.23 -> 0.94 ms, batch64 543 -> 474 ms against node records)
RECORDS = "auto"


def _layout(N, pair, dim):
    if RECORDS in ("cell", "auto"):
        return N.LAYOUT_CELL_PAIR if pair else N.LAYOUT_CELL
    return N.LAYOUT_PAIR if pair else N.LAYOUT_SINGLE


def _records_name(dim):
    return "cell" if RECORDS in ("cell", "auto") else "node"


def build_abc128(torch, n=128, ns=128, delta=1e-3, h=0.01, steps=1000):
    import numpy as np

    from pyftle_amd import _native as N
    from pyftle_amd.field import DeviceField, GridGeometry

    L = 2.0 * math.pi
    sp = L / n
    geom = GridGeometry.uniform((n, n, n), (0.0, 0.0, 0.0), (sp, sp, sp), periodic=(True, True, True))
    field = DeviceField(geom, dtype="float32", layout=_layout(N, False, 3), ring_slots=1, n_snapshots=1)
    ax = np.arange(n) * sp
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")  # flat ix + n*(iy + n*iz)
    A, B, C = math.sqrt(3.0), math.sqrt(2.0), 1.0
    vel = np.stack([A * np.sin(Z) + C * np.cos(Y), B * np.sin(X) + A * np.cos(Z),
                    C * np.sin(Y) + B * np.cos(X)]).reshape(3, -1)
    field.pack(0, vel)
    dev = torch.device("cuda")
    s = (torch.arange(ns, device=dev, dtype=torch.float64) + 0.5) * (L / ns)
    cz, cy, cx = torch.meshgrid(s, s, s, indexing="ij")
    centers = torch.stack([cx.reshape(-1), cy.reshape(-1), cz.reshape(-1)], dim=1)
    blocks = []
    for a in range(3):
        e = torch.zeros(3, dtype=torch.float64, device=dev)
        e[a] = delta
        blocks += [centers - e, centers + e] if a == 0 else [centers + e, centers - e]
    pos = torch.cat(blocks, dim=0).contiguous()
    nseeds = ns ** 3
    bytes_per_seed_step = 6 * 4 * 1 * 3 * 8 * 4  # nb * stages * snapshots * dim * 2^dim * 4 B
    return dict(name="abc128", field=field, pos=pos, n_seeds=nseeds, steps=steps, h=h,
                time_mode="steady", t_start=0.0, map_period=steps * h,
                bytes_per_seed_step=bytes_per_seed_step, dim=3,
                host=dict(vel=vel, n=n, sp=sp, periodic=True),
                config={"workload": "abc128", "grid": [n, n, n], "periodic": True,
                        "seeds": [ns, ns, ns], "neighbors": 6, "delta": delta, "integrator": "rk4",
                        "h": h, "steps": steps, "interp": "trilinear", "field_dtype": "f32"})


def _double_gyre(t, X, Y, A=0.1, eps=0.25, omega=2.0 * math.pi / 10.0):
    import numpy as np

    s = math.sin(omega * t)
    f = eps * s * X**2 + (1.0 - 2.0 * eps * s) * X
    dfdx = 2.0 * eps * s * X + (1.0 - 2.0 * eps * s)
    u = -math.pi * A * np.sin(math.pi * f) * np.cos(math.pi * Y)
    v = math.pi * A * np.cos(math.pi * f) * np.sin(math.pi * Y) * dfdx
    return u, v


def build_gyre(torch, nx=512, ny=256, n_snap=101, nsx=400, nsy=200, delta=1e-3, h=0.05, steps=200):
    import numpy as np

    from pyftle_amd import _native as N
    from pyftle_amd.field import DeviceField, GridGeometry

    geom = GridGeometry.uniform((nx, ny), (0.0, 0.0), (2.0 / (nx - 1), 1.0 / (ny - 1)))
    field = DeviceField(geom, dtype="float32", layout=_layout(N, True, 2), ring_slots=n_snap,
                       n_snapshots=n_snap, t0=0.0, dt_snap=0.1)
    x = np.linspace(0.0, 2.0, nx)
    y = np.linspace(0.0, 1.0, ny)
    X, Y = np.meshgrid(x, y)  # flat ix + nx*iy
    snaps = []
    for k in range(n_snap):
        u, v = _double_gyre(0.1 * k, X, Y)
        vel = np.stack([u.ravel(), v.ravel()])
        field.pack(k, vel)
        snaps.append(vel)
    dev = torch.device("cuda")
    xs = (torch.arange(nsx, device=dev, dtype=torch.float64) + 0.5) * (2.0 / nsx)
    ys = (torch.arange(nsy, device=dev, dtype=torch.float64) + 0.5) * (1.0 / nsy)
    cy, cx = torch.meshgrid(ys, xs, indexing="ij")
    centers = torch.stack([cx.reshape(-1), cy.reshape(-1)], dim=1)
    ex = torch.tensor([delta, 0.0], dtype=torch.float64, device=dev)
    ey = torch.tensor([0.0, delta], dtype=torch.float64, device=dev)
    pos = torch.cat([centers - ex, centers + ex, centers + ey, centers - ey], dim=0).contiguous()
    bytes_per_seed_step = 4 * 4 * 2 * 2 * 4 * 4
    return dict(name="gyre512", field=field, pos=pos, n_seeds=nsx * nsy, steps=steps, h=h,
                time_mode="linear", t_start=0.0, map_period=steps * h,
                bytes_per_seed_step=bytes_per_seed_step, dim=2,
                host=dict(snaps=snaps, nx=nx, ny=ny),
                config={"workload": "gyre512", "grid": [nx, ny], "snapshots": n_snap,
                        "seeds": [nsx, nsy], "neighbors": 4, "delta": delta, "integrator": "rk4",
                        "h": h, "steps": steps, "interp": "bilinear+linear-time", "field_dtype": "f32"})


def batch64_table(n_fields=64, seed=0):
    """The sampled batch: eps~U[0.1,0.3], omega~U[0.5,1.5]*2pi/10, grid, seed grid
    and window drawn with numpy.random.default_rng(seed) (committed as
    tests/golden/batch64_table.json)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    grids = [(128, 64), (256, 128), (512, 256), (1024, 512)]
    seedg = [(64, 32), (128, 64), (256, 128), (512, 256), (1024, 512), (2048, 1024)]
    Ts = [5.0, 10.0, 15.0, 20.0]
    out = []
    for _ in range(n_fields):
        eps = float(rng.uniform(0.1, 0.3))
        omega = float(rng.uniform(0.5, 1.5) * 2.0 * math.pi / 10.0)
        g = grids[int(rng.integers(len(grids)))]
        sg = seedg[int(rng.integers(len(seedg)))]
        T = Ts[int(rng.integers(len(Ts)))]
        out.append({"eps": eps, "omega": omega, "grid": list(g), "seeds": list(sg), "T": T})
    return out


def _gyre_field_on_device(torch, nx, ny, n_snap, eps, omega, dt_snap=0.1, A=0.1, dtype="float32"):
    from pyftle_amd import _native as N
    from pyftle_amd.field import DeviceField, GridGeometry

    geom = GridGeometry.uniform((nx, ny), (0.0, 0.0), (2.0 / (nx - 1), 1.0 / (ny - 1)))
    field = DeviceField(geom, dtype=dtype, layout=_layout(N, True, 2) if dtype == "float32" else N.LAYOUT_PAIR,
                        ring_slots=n_snap, n_snapshots=n_snap, t0=0.0, dt_snap=dt_snap)
    dev = torch.device("cuda")
    x = torch.linspace(0.0, 2.0, nx, dtype=torch.float64, device=dev)
    y = torch.linspace(0.0, 1.0, ny, dtype=torch.float64, device=dev)
    Y, X = torch.meshgrid(y, x, indexing="ij")  # flat ix + nx*iy
    # every snapshot in one batch of element-wise kernels (a few dozen dispatches per field,
    # not a few dozen per snapshot), then one pack per snapshot; chunks bound the memory
    sn_all = torch.tensor([math.sin(omega * dt_snap * k) for k in range(n_snap)], dtype=torch.float64, device=dev)
    chunk = max(1, (1 << 26) // (nx * ny))
    for k0 in range(0, n_snap, chunk):
        sn = sn_all[k0:k0 + chunk][:, None, None]
        f = (eps * sn) * X * X + (1.0 - 2.0 * eps * sn) * X
        dfdx = (2.0 * eps * sn) * X + (1.0 - 2.0 * eps * sn)
        u = -math.pi * A * torch.sin(math.pi * f) * torch.cos(math.pi * Y)
        v = math.pi * A * torch.cos(math.pi * f) * torch.sin(math.pi * Y) * dfdx
        uv = torch.stack([u.reshape(u.shape[0], -1), v.reshape(v.shape[0], -1)], dim=1)  # (chunk, 2, n)
        for j in range(uv.shape[0]):
            field.pack(k0 + j, uv[j])
        del f, dfdx, u, v, uv
    return field


def _gyre_seeds(torch, nsx, nsy, delta=1e-3):
    dev = torch.device("cuda")
    xs = (torch.arange(nsx, device=dev, dtype=torch.float64) + 0.5) * (2.0 / nsx)
    ys = (torch.arange(nsy, device=dev, dtype=torch.float64) + 0.5) * (1.0 / nsy)
    cy, cx = torch.meshgrid(ys, xs, indexing="ij")
    c = torch.stack([cx.reshape(-1), cy.reshape(-1)], dim=1)
    ex = torch.tensor([delta, 0.0], dtype=torch.float64, device=dev)
    ey = torch.tensor([0.0, delta], dtype=torch.float64, device=dev)
    return torch.cat([c - ex, c + ex, c + ey, c - ey], dim=0).contiguous()


def build_batch64(torch, h=0.05, n_fields=64):
    from pyftle_amd.batch import Window

    table = batch64_table(n_fields)
    windows = []
    for row in table:
        nx, ny = row["grid"]
        n_snap = int(round(row["T"] / 0.1)) + 1
        field = _gyre_field_on_device(torch, nx, ny, n_snap, row["eps"], row["omega"])
        pos = _gyre_seeds(torch, *row["seeds"])
        steps = int(round(row["T"] / h))
        windows.append(Window(field=field, positions=pos, n_steps=steps, h=h, map_period=steps * h))
    n_seeds = sum(w.positions.shape[0] // 4 for w in windows)
    seed_steps = sum(w.positions.shape[0] // 4 * w.n_steps for w in windows)
    return dict(name="batch64", windows=windows, n_seeds=n_seeds, seed_steps=seed_steps, dim=2,
                bytes_per_seed_step=1024, table=table,
                config={"workload": "batch64", "fields": n_fields, "integrator": "rk4", "h": h,
                        "interp": "bilinear+linear-time", "field_dtype": "f32",
                        "table": "tests/golden/batch64_table.json", "seeds_total": n_seeds,
                        "seed_steps_total": seed_steps})


def build_surrogate2d(torch, nsx=2048, nsy=1024, n_t0=32, members=4, epochs=200, n_traj=4_000_000, depth=5):
    """Double-gyre flow-map surrogate: 512x256 snapshots over t in [0, 20] (the flow is
    10-periodic, windows start in [0, 10)), RK4 h=0.05 over T=10."""
    from pyftle_amd import surrogate_ftle as SF

    field = _gyre_field_on_device(torch, 512, 256, 201, 0.25, 2.0 * math.pi / 10.0)
    return dict(name="surrogate2d", field=field, nsx=nsx, nsy=nsy, n_t0=n_t0, members=members,
                epochs=epochs, n_traj=n_traj, T=10.0, h=0.05, dim=2, SF=SF, depth=depth,
                config={"workload": "surrogate2d", "flow": "double gyre A=0.1 eps=0.25 omega=2pi/10",
                        "grid": [512, 256], "snapshots": 201, "seeds": [nsx, nsy], "t0_values": n_t0,
                        "T": 10.0, "rk4_h": 0.05,
                        "mlp": f"64 Fourier freqs on (x,y,t0), {depth}x256 SiLU -> 2",
                        "members": members, "epochs": epochs, "trajectories": n_traj,
                        "fallback_tol": 1e-3, "dtype": "bf16 MFMA, fp32 accumulate"})


def surrogate2d_measure(torch, wl, args):
    """Surrogate FTLE fields (ensemble + RK4 fallback) against plain GPU RK4 on the same
    seed-windows, both timed with HIP events on the current stream.  Returns the summary
    object of the bench line (and the batched RK4 plan, to be closed by the caller)."""
    from pyftle_amd import surrogate as S

    SF = wl["SF"]
    field = wl["field"]
    t = time.perf_counter()
    X, D = SF.make_training_data(field, wl["n_traj"], wl["T"], wl["h"], (0.0, 10.0), (0.0, 0.0), (2.0, 1.0),
                                 n_bins=500)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t
    spec = S.MLPSpec(sigma=1.0, depth=wl.get("depth", 5))
    t = time.perf_counter()
    ens = SF.fit_ensemble(X, D, spec, members=wl["members"], epochs=wl["epochs"], batch=16384, lr=3e-3, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t
    del X, D
    nsx, nsy = wl["nsx"], wl["nsy"]
    dev = torch.device("cuda")
    xs = (torch.arange(nsx, device=dev, dtype=torch.float64) + 0.5) * (2.0 / nsx)
    ys = (torch.arange(nsy, device=dev, dtype=torch.float64) + 0.5) * (1.0 / nsy)
    cy, cx = torch.meshgrid(ys, xs, indexing="ij")
    seeds = torch.stack([cx.reshape(-1), cy.reshape(-1)], 1).contiguous()
    t0s = [10.0 * k / wl["n_t0"] for k in range(wl["n_t0"])]
    n_sw = seeds.shape[0] * len(t0s)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            out = fn()
        b.record()
        torch.cuda.synchronize()
        return out, a.elapsed_time(b) / reps

    reps = max(1, min(3, args.steps))  # timed repetitions (the work is ~1 s per pass)
    plan, wins = SF.rk4_ftle_windows(field, seeds, t0s, wl["T"], wl["h"])
    _, rk4_ms = timed(lambda: plan.run(), reps)
    log(f"plain RK4: {rk4_ms:.1f} ms for {n_sw} seed-windows")
    fpq = S.mlp_flops_per_query(spec)
    M = wl["members"]

    def rmse_of(fields):
        rm = [SF.relative_rmse(f, w.ftle) for f, w in zip(fields, wins)]
        return sum(rm) / len(rm), max(rm)

    # surrogate alone: the FTLE gate's pass (every member's Jacobian, rows = M (1 + dim)) and the
    # displacement gate's cheapest pass (one member's Jacobian, rows = 1 + dim + M - 1)
    alone = {}
    for gate, jm in (("ftle", M), ("displacement", 1)):
        sf = SF.SurrogateFTLE(ens, field, wl["T"], wl["h"], gate=gate, jacobian_members=jm)
        preds, ms = timed(lambda: [sf.predict(seeds, t0) for t0 in t0s], reps)
        rows = n_sw * (jm * 3 + (M - jm))
        mean, worst = rmse_of([p[0] for p in preds])
        sp = torch.cat([p[1] for p in preds])
        sub = sp[torch.randperm(sp.numel(), device=sp.device)[:1 << 20]].float()
        qs = torch.quantile(sub, torch.tensor([0.5, 0.9, 0.99], device=sp.device)).tolist()
        alone[gate] = {"ms": ms, "rmse_mean": mean, "rmse_max": worst, "jacobian_members": jm,
                       "tflops": rows * fpq / (ms * 1e-3) / 1e12, "rows": rows,
                       "speedup_vs_gpu_rk4": rk4_ms / ms, "disagreement_quantiles_50_90_99": qs}
        log(f"surrogate alone ({gate} gate pass): {ms:.1f} ms, FTLE rmse/range {mean:.4f}, "
            f"disagreement p50/p90/p99 {qs}")
        del preds, sp
    # composite fields: RK4 wherever the disagreement exceeds tol (the fallback seeds of all
    # window starts in one batched launch); times include the probe, the ensemble and the
    # fallback.  The probe plan (the default) sends a tile to RK4 whole when more than half
    # of its probed seeds fail.
    # The last sweep gates on the first two members only (the ensemble's first two, as trained:
    # a cheaper setting of the same models; one Jacobian member plus one primal row per seed).
    sweeps = {}
    configs = [("ftle", "ftle", M, M, (1e-3, 3e-3, 1e-2, 3e-2, 1e-1)),
               ("displacement", "displacement", 1, M, (1e-3, 1e-2, 3e-2, 1e-1))]
    if M > 2:
        configs.append(("displacement_2_members", "displacement", 1, 2, (1e-2, 3e-2, 1e-1)))
    for key, gate, jm, n_mem, tols in configs:
        sub = ens if n_mem == M else S.FlowMapEnsemble(ens.members[:n_mem])
        sf = SF.SurrogateFTLE(sub, field, wl["T"], wl["h"], gate=gate, jacobian_members=jm)
        share = sf.tile_fallback_max
        rows = []
        for tol in tols:
            sf.tol = tol
            res, ms = timed(lambda: sf.ftle_windows(seeds, t0s), 1)
            mean, worst = rmse_of([r.ftle for r in res])
            fb = sum(int(r.fallback.sum().item()) for r in res) / n_sw
            skipped = sum(int(torch.isnan(r.spread).sum().item()) for r in res) / n_sw
            rows.append({"tol": tol, "ms": ms, "seeds_per_s": n_sw / (ms * 1e-3), "fallback_fraction": fb,
                         "tile_rk4_fraction": skipped, "tile_fallback_max": share, "members": n_mem,
                         "rmse_mean": mean, "rmse_max": worst, "speedup_vs_gpu_rk4": rk4_ms / ms})
            log(f"{gate} gate ({n_mem} members) tol {tol:g}: {ms:.1f} ms, fallback {fb:.4f} (tiles to RK4 "
                f"{skipped:.4f}), rmse/range {mean:.4f}")
            del res
        sweeps[key] = rows
    # the specified setting without the probe plan (every seed through the ensemble)
    sf = SF.SurrogateFTLE(ens, field, wl["T"], wl["h"], gate="ftle", tol=1e-3, probe=0)
    res, ms = timed(lambda: sf.ftle_windows(seeds, t0s), 1)
    mean, worst = rmse_of([r.ftle for r in res])
    fb = sum(int(r.fallback.sum().item()) for r in res) / n_sw
    unplanned = {"tol": 1e-3, "gate": "ftle", "ms": ms, "fallback_fraction": fb, "rmse_max": worst,
                 "speedup_vs_gpu_rk4": rk4_ms / ms}
    log(f"ftle gate tol 1e-3 without the probe plan: {ms:.1f} ms, fallback {fb:.4f}")
    del res
    spec_row = next(r for r in sweeps["ftle"] if r["tol"] == 1e-3)
    ok = [r for g in sweeps.values() for r in g if r["rmse_max"] <= 0.02]
    best = max(ok, key=lambda r: r["speedup_vs_gpu_rk4"]) if ok else None
    best_gate = next((g for g, rr in sweeps.items() if best in rr), None)
    best_gate = "displacement" if best_gate == "displacement_2_members" else best_gate
    # training pays for itself once the time saved per field set has covered it
    saved = rk4_ms - (best["ms"] if best else spec_row["ms"])
    a = alone["ftle"]
    return {
        "workload": "surrogate2d", "seed_windows": n_sw, "members": M, "network": "64 Fourier freqs, 5x256 SiLU",
        "gate": "ftle: max over members |FTLE_member - FTLE_ensemble| / (FTLE range of the window's probe)",
        "specified": {"tol": 1e-3, "gate": "ftle", "ms": spec_row["ms"], "fallback_fraction": spec_row["fallback_fraction"],
                      "tile_rk4_fraction": spec_row["tile_rk4_fraction"],
                      "rmse_max": spec_row["rmse_max"], "speedup_vs_gpu_rk4": spec_row["speedup_vs_gpu_rk4"],
                      "seeds_per_s": spec_row["seeds_per_s"], "without_probe_plan": unplanned},
        "best_within_2pct": None if best is None else dict(best, gate=best_gate),
        "surrogate_alone": alone, "tolerance_sweeps": sweeps,
        "mfma": {"achieved": a["tflops"], "peak": 2500.0, "unit": "TFLOP/s", "frac": a["tflops"] / 2500.0,
                 "kernel": "fused_mlp_kernel<256,6,2,2> (every member with its 2 tangents)",
                 "flops_per_query": fpq, "rows": a["rows"]},
        "rk4": {"ms_per_field_set": rk4_ms, "seeds_per_s": n_sw / (rk4_ms * 1e-3),
                "seed_steps_per_s": n_sw * round(wl["T"] / wl["h"]) / (rk4_ms * 1e-3)},
        "train_s": t_train, "data_s": t_data,
        "break_even_field_sets": (t_train + t_data) / (saved * 1e-3) if saved > 0 else None,
    }, plan


def run_surrogate(torch, wl, args):
    """The 2D surrogate as its own bench line: value = seeds/s of the specified setting."""
    summ, plan = surrogate2d_measure(torch, wl, args)
    spec = summ["specified"]
    result = {
        "metric": METRIC, "value": spec["seeds_per_s"], "unit": "seeds/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": spec["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": dict(wl["config"], fallback_tol=1e-3, gate="ftle"),
        "surrogate": summ,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(_gyre_cpu_workload(torch, wl), budget_s=15.0),
    }
    print(json.dumps(result), flush=True)
    plan.close()


def _abc_unsteady_field_on_device(torch, n, t_first, n_slots, dt_snap, cell=False):
    """Unsteady ABC flow, A(t) = sqrt(3) + 0.5 sin(pi t / 5), B = sqrt(2), C = 1, on a
    periodic n^3 grid over [0, 2 pi)^3 (no endpoint), FP32 time-pair snapshots (node or
    cell records) at t_first + k dt_snap, k < n_slots."""
    from pyftle_amd import _native as N
    from pyftle_amd.field import DeviceField, GridGeometry

    L = 2.0 * math.pi
    geom = GridGeometry.uniform((n, n, n), (0.0,) * 3, (L / n,) * 3, periodic=(True,) * 3)
    field = DeviceField(geom, dtype="float32", layout=N.LAYOUT_CELL_PAIR if cell else N.LAYOUT_PAIR,
                        ring_slots=n_slots, n_snapshots=n_slots, t0=t_first, dt_snap=dt_snap)
    ax = torch.arange(n, dtype=torch.float64, device="cuda") * (L / n)
    Z, Y, X = torch.meshgrid(ax, ax, ax, indexing="ij")  # flat ix + n*(iy + n*iz)
    B, C = math.sqrt(2.0), 1.0
    sx, cx, sy, cy, sz, cz = X.sin(), X.cos(), Y.sin(), Y.cos(), Z.sin(), Z.cos()
    del X, Y, Z
    for k in range(n_slots):
        a = math.sqrt(3.0) + 0.5 * math.sin(math.pi * (t_first + k * dt_snap) / 5.0)
        field.pack(k, torch.stack([a * sz + C * cy, B * sx + a * cz, C * sy + B * cx]).reshape(3, -1))
    torch.cuda.synchronize()
    return field


def _abc_cpu_workload(torch, wl, seeds, delta=1e-3):
    """The 3D surrogate's flow as the reference CPU path sees it: the snapshot at the
    first window start, frozen, on the host, and the 6-neighbour groups of the seeds."""
    import numpy as np

    n = wl["config"]["grid"][0]
    L = 2.0 * math.pi
    ax = np.arange(n) * (L / n)
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    a = math.sqrt(3.0) + 0.5 * math.sin(math.pi * wl["t0s"][0] / 5.0)
    B, C = math.sqrt(2.0), 1.0
    vel = np.stack([a * np.sin(Z) + C * np.cos(Y), B * np.sin(X) + a * np.cos(Z),
                    C * np.sin(Y) + B * np.cos(X)]).reshape(3, -1)
    blocks = []
    for c in range(3):
        e = torch.zeros(3, dtype=torch.float64, device=seeds.device)
        e[c] = delta
        blocks += [seeds - e, seeds + e] if c == 0 else [seeds + e, seeds - e]
    return dict(name="surrogate3d", pos=torch.cat(blocks, 0), n_seeds=seeds.shape[0], dim=3, h=wl["h"],
                steps=int(round(wl["T"] / wl["h"])), host=dict(vel=vel, n=n, sp=L / n))


def build_surrogate3d(torch, n=192, ns=256, t0s=(0.0, 2.5, 5.0, 7.5), members=4, epochs=100, n_traj=4_000_000,
                      depth=6):
    """Unsteady ABC on a periodic 192^3 grid: 64 snapshots per period 10 (dt = 10/64),
    stored from t = -1.25 to 16.25 so that training windows start in [-1.25, 11.25)
    and the 4 inference windows (T = 5, RK4 h = 0.01) lie inside."""
    from pyftle_amd import surrogate_ftle as SF

    dts = 10.0 / 64.0
    field = _abc_unsteady_field_on_device(torch, n, -1.25, 113, dts)
    return dict(name="surrogate3d", field=field, ns=ns, t0s=list(t0s), members=members, epochs=epochs,
                n_traj=n_traj, T=5.0, h=0.01, dim=3, SF=SF, depth=depth,
                config={"workload": "surrogate3d", "flow": "ABC, A(t)=sqrt3+0.5 sin(pi t/5), B=sqrt2, C=1",
                        "grid": [n, n, n], "periodic": True, "snapshots_per_period": 64,
                        "stored_time": [-1.25, 16.25], "seeds": [ns, ns, ns], "neighbors": 6, "delta": 1e-3,
                        "t0_values": list(t0s), "T": 5.0, "rk4_h": 0.01,
                        "mlp": f"64 Fourier freqs (sigma 0.5) on (x,y,z,t0), {depth}x512 SiLU -> 3",
                        "members": members, "epochs": epochs, "trajectories": n_traj,
                        "accuracy_reference": "RK4 on a 1% held-out seed subset (rng seed 1)",
                        "dtype": "bf16 MFMA, fp32 accumulate"})


def surrogate3d_measure(torch, wl, args, full=True):
    """3D surrogate FTLE fields against plain GPU RK4 on all 256^3 seeds of the first window
    (times with HIP events; accuracy against the full RK4 field): the specified gate, the
    surrogate alone, and the gated settings end to end.  ``full`` adds the node-record RK4,
    the specified gate without the probe plan, the held-out 1% of all four windows and more
    tolerances.  Returns the summary object and the RK4 figures."""
    from pyftle_amd import surrogate as S

    SF = wl["SF"]
    field = wl["field"]
    L = 2.0 * math.pi
    t0w = wl["t0s"][0]
    t = time.perf_counter()
    X, D = SF.make_training_data(field, wl["n_traj"], wl["T"], wl["h"], (-1.25, 11.25), (0.0,) * 3, (L,) * 3,
                                 n_bins=400)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t
    spec = S.MLPSpec(d_in=4, d_out=3, n_freq=64, width=512, depth=wl.get("depth", 6), sigma=0.5,
                     scale=(0.5, 0.5, 0.5, 0.1), out_scale=2.0)
    t = time.perf_counter()
    ens = SF.fit_ensemble(X, D, spec, members=wl["members"], epochs=wl["epochs"], batch=16384, lr=2e-3, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t
    del X, D
    ns = wl["ns"]
    dev = torch.device("cuda")
    sax = (torch.arange(ns, device=dev, dtype=torch.float64) + 0.5) * (L / ns)
    cz, cy, cx = torch.meshgrid(sax, sax, sax, indexing="ij")
    seeds = torch.stack([cx.reshape(-1), cy.reshape(-1), cz.reshape(-1)], 1).contiguous()
    del cx, cy, cz
    sf = SF.SurrogateFTLE(ens, field, wl["T"], wl["h"], chunk=1 << 19)

    def timed(fn, reps=1):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            out = fn()
        b.record()
        torch.cuda.synchronize()
        return out, a.elapsed_time(b) / reps

    n_seeds = seeds.shape[0]
    rk4 = {}
    if full:
        _, rk4["ms_node_records"] = timed(lambda: sf.rk4_ftle(seeds, t0w))
    # the fastest plain RK4 of this build: cell-pair records, in a ring holding the snapshots
    # of the timed window only (the whole stored range would take 160 GB as cell pairs)
    dts, t_first = 10.0 / 64.0, -1.25
    k0 = int(math.floor((t0w - t_first) / dts))
    n_win = int(math.ceil(wl["T"] / dts)) + 2
    cell_field = _abc_unsteady_field_on_device(torch, wl["config"]["grid"][0], t_first + k0 * dts, n_win, dts,
                                               cell=True)
    sf_cell = SF.SurrogateFTLE(ens, cell_field, wl["T"], wl["h"], chunk=1 << 19)
    ref_cell, rk4_ms = timed(lambda: sf_cell.rk4_ftle(seeds, t0w))
    rk4.update({"ms_all_seeds_one_window": rk4_ms, "seeds_per_s": n_seeds / (rk4_ms * 1e-3),
                "records": "cell pairs (ring of the timed window's snapshots)"})
    if full:
        ref_node = sf.rk4_ftle(seeds, t0w)
        # the two rings start at different times, so the time weights could round differently
        rk4["max_ftle_diff_over_range_node_vs_cell"] = \
            ((ref_cell - ref_node).abs().max() / (ref_node.max() - ref_node.min())).item()
        del ref_node
    del sf_cell
    torch.cuda.empty_cache()
    log(f"plain RK4, all {n_seeds} seeds, one window: {rk4_ms:.1f} ms with cell-pair records")
    # the surrogate alone on every seed of the window, against the full RK4 field: every member's
    # Jacobian (the FTLE gate's pass) and one member's (the displacement gate's pass)
    sf_ftle = SF.SurrogateFTLE(ens, field, wl["T"], wl["h"], gate="ftle", chunk=1 << 19)
    sf_disp = SF.SurrogateFTLE(ens, field, wl["T"], wl["h"], gate="displacement", jacobian_members=1,
                               chunk=1 << 19)
    (f4, _), jm4 = timed(lambda: sf_ftle.predict(seeds, t0w))
    rmse4 = SF.relative_rmse(f4, ref_cell)
    del f4
    (f1, dsp), jm1 = timed(lambda: sf_disp.predict(seeds, t0w))
    rmse1 = SF.relative_rmse(f1, ref_cell)
    sub = dsp[torch.randperm(dsp.numel(), device=dev)[:1 << 20]].float()
    disp_q = torch.quantile(sub, torch.tensor([0.5, 0.9, 0.99], device=dev)).tolist()
    del f1, dsp, sub
    log(f"surrogate alone: every member's Jacobian {jm4:.1f} ms (rmse/range {rmse4:.4f}), one member's "
        f"{jm1:.1f} ms (rmse/range {rmse1:.4f}); displacement disagreement p50/p90/p99 {disp_q}")
    fpq = S.mlp_flops_per_query(spec)
    rows4 = n_seeds * wl["members"] * 4
    # the specified setting end to end: FTLE gate at 1e-3, the fallback on the cell-pair ring;
    # the probe plan (the default) sends a tile to RK4 whole when more than half of its probed
    # seeds fail
    spec_runs = {}
    for name, probe in (("planned", 64), ("without_probe_plan", 0)) if full else (("planned", 64),):
        sf_spec = SF.SurrogateFTLE(ens, cell_field, wl["T"], wl["h"], gate="ftle", tol=1e-3, chunk=1 << 19,
                                   probe=probe)
        res, ms = timed(lambda: sf_spec.ftle(seeds, t0w))
        spec_runs[name] = {"ms": ms, "fallback_fraction": res.fallback.float().mean().item(),
                           "tile_rk4_fraction": torch.isnan(res.spread).float().mean().item(),
                           "speedup_vs_gpu_rk4": rk4_ms / ms}
        log(f"specified setting ({name}): {ms:.1f} ms, fallback {spec_runs[name]['fallback_fraction']:.4f}")
        del res
    share = sf_spec.tile_fallback_max
    # looser gates end to end on every seed of the same window, judged by the full RK4 field
    M = wl["members"]
    gates = ((("ftle", 3e-2, M), ("ftle", 1e-1, M), ("displacement", 0.1, M), ("displacement", 0.3, M),
              ("displacement", 1.0, M))
             if full else (("ftle", 1e-1, M), ("displacement", 0.3, M), ("displacement", 1.0, M)))
    if M > 2:  # gating on the ensemble's first two members only: a cheaper setting of the same models
        gates += (("displacement", 0.3, 2), ("displacement", 1.0, 2))
    measured = []
    for gate, tol, n_mem in gates:
        sub = ens if n_mem == M else S.FlowMapEnsemble(ens.members[:n_mem])
        sf_spec = SF.SurrogateFTLE(sub, cell_field, wl["T"], wl["h"], gate=gate, tol=tol, chunk=1 << 19,
                                   jacobian_members=1 if gate == "displacement" else None)
        res, ms = timed(lambda: sf_spec.ftle(seeds, t0w))
        row = {"gate": gate, "tol": tol, "members": n_mem, "t0": t0w, "ms": ms,
               "fallback_fraction": res.fallback.float().mean().item(),
               "tile_rk4_fraction": torch.isnan(res.spread).float().mean().item(),
               "rmse": SF.relative_rmse(res.ftle, ref_cell), "speedup_vs_gpu_rk4": rk4_ms / ms}
        measured.append(row)
        log(f"{gate} gate ({n_mem} members) tol {tol:g} on all seeds: {ms:.1f} ms, "
            f"fallback {row['fallback_fraction']:.4f}, "
            f"rmse/range {row['rmse']:.4f} against the full RK4 field, {row['speedup_vs_gpu_rk4']:.2f}x")
        del res
    del cell_field, sf_spec, ref_cell
    torch.cuda.empty_cache()
    summary = {}
    if full:
        # the FTLE gate on the held-out 1% of the seeds of all four windows (node-record RK4):
        # time estimated as the gate's pass plus RK4 on the fallback share
        g = torch.Generator(device=dev).manual_seed(1)
        check_idx = torch.randperm(n_seeds, generator=g, device=dev)[: n_seeds // 100].sort().values
        check_seeds = seeds[check_idx].contiguous()
        tols = (1e-3, 1e-2, 3e-2, 1e-1)
        per_t0 = []
        for t0 in wl["t0s"]:
            ref = sf.rk4_ftle(check_seeds, t0)
            f_all, dis = sf_ftle.predict(seeds, t0)
            f_j, d_j = f_all[check_idx], dis[check_idx]
            row = {"t0": t0, "ftle_gate_rmse": SF.relative_rmse(f_j, ref)}
            for tol in tols:
                bad = d_j > tol
                row[f"tol{tol:g}"] = {"fallback": bad.float().mean().item(),
                                      "rmse": SF.relative_rmse(torch.where(bad, ref, f_j), ref)}
            del f_all, dis
            f1, _ = sf_disp.predict(seeds, t0)
            row["jac1_rmse"] = SF.relative_rmse(f1[check_idx], ref)
            del f1
            log(f"t0 {t0} (held-out 1%): " + ", ".join(f"{k} {v}" for k, v in row.items() if k != "t0"))
            per_t0.append(row)
        sweep = []
        for tol in tols:
            f = sum(r[f"tol{tol:g}"]["fallback"] for r in per_t0) / len(per_t0)
            est = jm4 + f * rk4_ms
            sweep.append({"tol": tol, "fallback_fraction": f,
                          "rmse_mean": sum(r[f"tol{tol:g}"]["rmse"] for r in per_t0) / len(per_t0),
                          "rmse_max": max(r[f"tol{tol:g}"]["rmse"] for r in per_t0), "ms_estimate": est,
                          "speedup_vs_gpu_rk4_estimate": rk4_ms / est})
        summary.update({"tolerance_sweep_held_out": sweep, "per_t0_held_out": per_t0})
    sur_ms, fb = spec_runs["planned"]["ms"], spec_runs["planned"]["fallback_fraction"]
    ok = [r for r in measured if r["rmse"] <= 0.02]
    best = max(ok, key=lambda r: r["speedup_vs_gpu_rk4"]) if ok else None
    saved = rk4_ms - (best["ms"] if best else sur_ms)
    specified = {"tol": 1e-3, "gate": "ftle", "ms": sur_ms, "fallback_fraction": fb,
                 "tile_rk4_fraction": spec_runs["planned"]["tile_rk4_fraction"], "tile_fallback_max": share,
                 "speedup_vs_gpu_rk4": rk4_ms / sur_ms, "seeds_per_s": n_seeds / (sur_ms * 1e-3)}
    if full:
        specified["without_probe_plan"] = spec_runs["without_probe_plan"]
    summary.update({
        "workload": "surrogate3d", "seeds": n_seeds, "window_t0": t0w, "members": wl["members"],
        "network": f"64 Fourier freqs on (x,y,z,t0), {spec.depth}x512 SiLU -> 3",
        "gate": "ftle: max over members |FTLE_member - FTLE_ensemble| / (FTLE range of the window's probe); "
                "displacement: max member deviation of the displacement, domain units (the domain is 2 pi)",
        "specified": specified, "best_within_2pct": best, "measured_all_seeds_first_window": measured,
        "displacement_disagreement_quantiles_50_90_99": disp_q,
        "surrogate_alone": {"ftle": {"ms": jm4, "rmse": rmse4, "jacobian_members": wl["members"],
                                     "speedup_vs_gpu_rk4": rk4_ms / jm4},
                            "displacement": {"ms": jm1, "rmse": rmse1, "jacobian_members": 1,
                                             "speedup_vs_gpu_rk4": rk4_ms / jm1}},
        "mfma": {"achieved": rows4 * fpq / (jm4 * 1e-3) / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
                 "frac": rows4 * fpq / (jm4 * 1e-3) / 1e12 / 2500.0, "flops_per_query": fpq,
                 "kernel": "fused_mlp_kernel<512,4,3,3> (every member with its 3 tangents)"},
        "rk4": rk4, "train_s": t_train, "data_s": t_data,
        "break_even_windows": (t_train + t_data) / (saved * 1e-3) if saved > 0 else None})
    return summary, seeds


def run_surrogate3d(torch, wl, args):
    """The 3D surrogate as its own bench line: value = seeds/s of the specified setting."""
    summ, seeds = surrogate3d_measure(torch, wl, args, full=True)
    sp = summ["specified"]
    result = {
        "metric": METRIC, "value": sp["seeds_per_s"], "unit": "seeds/s", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": sp["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "data": "synthetic", "config": dict(wl["config"], fallback_tol=1e-3, gate="ftle"),
        "surrogate": summ,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(_abc_cpu_workload(torch, wl, seeds), budget_s=15.0),
    }
    print(json.dumps(result), flush=True)


def build_surrogate_infer(torch, nsx=2048, nsy=1024, n_t0=32, members=4):
    """Inference half of surrogate2d: the same 67.1M seed-windows through four
    random-init members of the same architecture (no training), for profiling the
    fused MLP kernel."""
    wl = build_surrogate2d(torch, nsx=nsx, nsy=nsy, n_t0=n_t0, members=members, epochs=0, n_traj=0)
    wl["name"] = "surrogate_infer"
    wl["config"] = dict(wl["config"], workload="surrogate_infer", weights="random init (no training)")
    for k in ("epochs", "trajectories"):
        wl["config"].pop(k, None)
    return wl


def _gyre_cpu_workload(torch, wl):
    """The surrogate's flow as the reference CPU path sees it: snapshot 0 of the
    512x256 gyre on the host and the seed-neighbour groups of a 400x200 seed grid
    (one window, frozen field, h = 0.05, 200 steps)."""
    import numpy as np

    nx, ny = wl["config"]["grid"]
    X, Y = np.meshgrid(np.linspace(0.0, 2.0, nx), np.linspace(0.0, 1.0, ny))
    u, v = _double_gyre(0.0, X, Y)
    nsx, nsy, delta = 400, 200, 1e-3
    dev = torch.device("cuda")
    xs = (torch.arange(nsx, device=dev, dtype=torch.float64) + 0.5) * (2.0 / nsx)
    ys = (torch.arange(nsy, device=dev, dtype=torch.float64) + 0.5) * (1.0 / nsy)
    cy, cx = torch.meshgrid(ys, xs, indexing="ij")
    c = torch.stack([cx.reshape(-1), cy.reshape(-1)], dim=1)
    ex = torch.tensor([delta, 0.0], dtype=torch.float64, device=dev)
    ey = torch.tensor([0.0, delta], dtype=torch.float64, device=dev)
    pos = torch.cat([c - ex, c + ex, c + ey, c - ey], dim=0).contiguous()
    return dict(name="gyre512", pos=pos, n_seeds=nsx * nsy, dim=2, h=wl["h"], steps=int(round(wl["T"] / wl["h"])),
                host=dict(snaps=[np.stack([u.ravel(), v.ravel()])], nx=nx, ny=ny))


def run_surrogate_infer(torch, wl, args):
    """Ensemble FTLE (no fallback) over all seed-windows, and the fused kernel alone
    timed with HIP events on the current stream (the roofline entry)."""
    from pyftle_amd import surrogate as S

    SF = wl["SF"]
    spec = S.MLPSpec(sigma=1.0)
    ens = S.FlowMapEnsemble([S.FlowMapMLP(spec, seed=1000 + k) for k in range(wl["members"])])
    nsx, nsy = wl["nsx"], wl["nsy"]
    dev = torch.device("cuda")
    xs = (torch.arange(nsx, device=dev, dtype=torch.float64) + 0.5) * (2.0 / nsx)
    ys = (torch.arange(nsy, device=dev, dtype=torch.float64) + 0.5) * (1.0 / nsy)
    cy, cx = torch.meshgrid(ys, xs, indexing="ij")
    seeds = torch.stack([cx.reshape(-1), cy.reshape(-1)], 1).contiguous()
    t0s = [10.0 * k / wl["n_t0"] for k in range(wl["n_t0"])]
    n_sw = seeds.shape[0] * len(t0s)
    sf = SF.SurrogateFTLE(ens, wl["field"], wl["T"], wl["h"], tol=float("inf"), gate="displacement")
    fpq = S.mlp_flops_per_query(spec)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    steps = max(1, args.steps)
    out = {}
    for jm in (wl["members"], 1):
        sf.jacobian_members = jm
        ms = timed(lambda: [sf.predict(seeds, t0) for t0 in t0s], steps)
        rows = n_sw * (jm * 3 + (wl["members"] - jm))
        out[jm] = {"ms": ms, "seeds_per_s": n_sw / (ms * 1e-3), "tflops": rows * fpq / (ms * 1e-3) / 1e12}
        log(f"ensemble FTLE, Jacobian from {jm} member(s): {ms:.1f} ms, {out[jm]['tflops']:.0f} TFLOP/s")
    # the kernel alone: one member, 2 tangents, 2M seeds per launch
    m0 = ens.members[0]
    x = torch.empty(seeds.shape[0], 3, dtype=torch.float32, device=dev)
    x[:, :2] = seeds.float()
    x[:, 2] = 2.5
    Y = torch.empty(3, seeds.shape[0], 2, dtype=torch.float32, device=dev)
    kms = timed(lambda: m0.forward_fused(x, n_tangents=2, out=Y), 10 * steps)
    kflops = seeds.shape[0] * 3 * fpq
    Y0 = torch.empty(1, seeds.shape[0], 2, dtype=torch.float32, device=dev)
    kms0 = timed(lambda: m0.forward_fused(x, n_tangents=0, out=Y0), 10 * steps)
    ach = kflops / (kms * 1e-3) / 1e12
    main = out[wl["members"]]
    cpu = None if args.no_cpu else cpu_baseline(_gyre_cpu_workload(torch, wl), budget_s=15.0)
    result = {
        "metric": METRIC, "value": main["seeds_per_s"], "unit": "seeds/s", "n_gpus": 1, "steps": steps,
        "warmup": 1, "ms_per_step": main["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16", "data": "synthetic", "config": wl["config"],
        "roofline": {"bound": "mfma", "achieved": ach, "peak": 2500.0, "unit": "TFLOP/s", "frac": ach / 2500.0,
                     "traffic": None,
                     "kernel": "fused_mlp_kernel<256,6,2,2> (2M seeds, primal + 2 tangent rows)",
                     "flops_per_launch": kflops, "avg_launch_ms": kms},
        "fused_primal_only": {"kernel": "fused_mlp_kernel<256,8,0,2>", "avg_launch_ms": kms0,
                              "tflops": seeds.shape[0] * fpq / (kms0 * 1e-3) / 1e12},
        "ensemble": {str(k): v for k, v in out.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


def build_surrogate3d_infer(torch, ns=256, members=4):
    """Inference half of surrogate3d: one window's 256^3 seeds through four random-init
    6x512 members (no field, no training), for profiling the width-512 fused kernel."""
    return dict(name="surrogate3d_infer", ns=ns, members=members, T=5.0,
                config={"workload": "surrogate3d_infer", "seeds": [ns, ns, ns], "t0": 2.5,
                        "mlp": "64 Fourier freqs (sigma 0.5) on (x,y,z,t0), 6x512 SiLU -> 3",
                        "members": members, "weights": "random init (no training)",
                        "dtype": "bf16 MFMA, fp32 accumulate"})


def run_surrogate3d_infer(torch, wl, args):
    """Ensemble FTLE (no fallback) of one 256^3-seed window, and the width-512 fused kernel
    alone on 2^19 seeds, timed with HIP events on the current stream."""
    from pyftle_amd import surrogate as S
    from pyftle_amd import surrogate_ftle as SF

    spec = S.MLPSpec(d_in=4, d_out=3, n_freq=64, width=512, depth=6, sigma=0.5, scale=(0.5, 0.5, 0.5, 0.1),
                     out_scale=2.0)
    ens = S.FlowMapEnsemble([S.FlowMapMLP(spec, seed=1000 + k) for k in range(wl["members"])])
    ns, L = wl["ns"], 2.0 * math.pi
    dev = torch.device("cuda")
    sax = (torch.arange(ns, device=dev, dtype=torch.float64) + 0.5) * (L / ns)
    cz, cy, cx = torch.meshgrid(sax, sax, sax, indexing="ij")
    seeds = torch.stack([cx.reshape(-1), cy.reshape(-1), cz.reshape(-1)], 1).contiguous()
    del cx, cy, cz
    sf = SF.SurrogateFTLE(ens, None, wl["T"], 0.01, tol=float("inf"), chunk=1 << 19, gate="displacement")
    fpq = S.mlp_flops_per_query(spec)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    steps = max(1, args.steps)
    n = seeds.shape[0]
    out = {}
    for jm in (wl["members"], 1):
        sf.jacobian_members = jm
        ms = timed(lambda: sf.predict(seeds, 2.5), steps)
        rows = n * (jm * 4 + (wl["members"] - jm))
        out[jm] = {"ms": ms, "seeds_per_s": n / (ms * 1e-3), "tflops": rows * fpq / (ms * 1e-3) / 1e12}
        log(f"ensemble FTLE, Jacobian from {jm} member(s): {ms:.1f} ms, {out[jm]['tflops']:.0f} TFLOP/s")
    m0 = ens.members[0]
    M = 1 << 19
    x = torch.empty(M, 4, dtype=torch.float32, device=dev)
    x[:, :3] = seeds[:M].float()
    x[:, 3] = 2.5
    Y = torch.empty(4, M, 3, dtype=torch.float32, device=dev)
    kms = timed(lambda: m0.forward_fused(x, n_tangents=3, out=Y), 10 * steps)
    Y0 = torch.empty(1, M, 3, dtype=torch.float32, device=dev)
    kms0 = timed(lambda: m0.forward_fused(x, n_tangents=0, out=Y0), 10 * steps)
    kflops = M * 4 * fpq
    ach = kflops / (kms * 1e-3) / 1e12
    main = out[wl["members"]]
    result = {
        "metric": METRIC, "value": main["seeds_per_s"], "unit": "seeds/s", "n_gpus": 1, "steps": steps,
        "warmup": 1, "ms_per_step": main["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16", "data": "synthetic", "config": wl["config"],
        "roofline": {"bound": "mfma", "achieved": ach, "peak": 2500.0, "unit": "TFLOP/s", "frac": ach / 2500.0,
                     "traffic": None,
                     "kernel": "fused_mlp_kernel<512,4,3,3> (2^19 seeds, primal + 3 tangent rows)",
                     "flops_per_launch": kflops, "avg_launch_ms": kms},
        "fused_primal_only": {"kernel": "fused_mlp_kernel<512,4,0,3>", "avg_launch_ms": kms0,
                              "tflops": M * fpq / (kms0 * 1e-3) / 1e12},
        "ensemble": {str(k): v for k, v in out.items()},
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


def build_mlp_train(torch, batch=16384, shape="2d", n_rows=1 << 21):
    """One surrogate member trained on synthetic smooth targets, to time a training step:
    the 2D network (64 Fourier frequencies on (x, y, t0), 5 x 256 SiLU, 2 outputs) or the 3D
    one (on (x, y, z, t0), 6 x 512 SiLU, 3 outputs)."""
    from pyftle_amd import surrogate as S

    g = torch.Generator(device="cuda").manual_seed(0)
    if shape == "3d":
        X = torch.rand(n_rows, 4, generator=g, device="cuda") * torch.tensor([6.28, 6.28, 6.28, 10.0], device="cuda")
        D = torch.stack([0.5 * torch.sin(X[:, 2] + 0.2 * X[:, 3]) + 0.3 * torch.cos(X[:, 1]),
                         0.5 * torch.cos(X[:, 2]) + 0.4 * torch.sin(X[:, 0]),
                         0.3 * torch.sin(X[:, 1]) + 0.4 * torch.cos(X[:, 0] - 0.1 * X[:, 3])], dim=1)
        spec = S.MLPSpec(d_in=4, d_out=3, n_freq=64, width=512, depth=6, sigma=0.5, scale=(0.5, 0.5, 0.5, 0.1),
                         out_scale=2.0)
        mlp = "64 Fourier freqs on (x,y,z,t0), 6x512 SiLU -> 3"
    else:
        X = torch.rand(n_rows, 3, generator=g, device="cuda") * torch.tensor([2.0, 1.0, 10.0], device="cuda")
        D = torch.stack([0.1 * torch.sin(3.0 * X[:, 0] + X[:, 2]) * torch.cos(2.0 * X[:, 1]),
                         0.1 * torch.cos(2.0 * X[:, 0]) * torch.sin(3.0 * X[:, 1] - 0.5 * X[:, 2])], dim=1)
        spec = S.MLPSpec(sigma=1.0, depth=5, width=256)
        mlp = "64 Fourier freqs on (x,y,t0), 5x256 SiLU -> 2"
    return dict(name="mlp_train", X=X, D=D.contiguous(), spec=spec, batch=batch, S=S,
                config={"workload": "mlp_train", "mlp": mlp, "batch": batch, "rows": n_rows,
                        "data": "synthetic smooth targets"})


def run_tile_occupancy(torch, wl, args):
    """Ceiling of an LDS tile path for the advection kernel: the share of particles that are
    still inside their work-group's tile -- the box of cells (as many as one work-group's
    share of LDS holds: 80 KB at two work-groups per CU) centred on the work-group's initial
    cell -- as the window advances.  Only samples in the tile could be served from LDS
    instead of the texture path, so this fraction bounds the TD time such a path could save.
    Work-groups are the bench's launch: the tuned number of seed groups per work-group in
    Morton order.  The window runs in chunks of steps through pyftle_advect; positions are
    compared with the tiles between chunks."""
    from pyftle_amd import window as W
    from pyftle_amd.tuning import best_config, shape_class

    field, dim = wl["field"], wl["dim"]
    nb = 2 * dim
    n = wl["n_seeds"]
    cfg = best_config(shape_class(dim, n, wl["time_mode"], records=_records_name(dim)))
    G = cfg.groups_per_block
    rec = 96 if dim == 3 else 64  # bytes per cell record (3D single level, 2D time pair)
    side = int((80 * 1024 // rec) ** (1.0 / dim))
    order = W.seed_order(field, wl["pos"], nb).long()
    geo = field.geom
    org = torch.tensor(geo.origin[:dim], dtype=torch.float64, device="cuda")
    spc = torch.tensor(geo.spacing[:dim], dtype=torch.float64, device="cuda")
    per = torch.tensor([bool(p) for p in (geo.periodic or (False,) * dim)][:dim], device="cuda")
    ncell = torch.tensor(geo.n[:dim], dtype=torch.float64, device="cuda")

    def cells(p):
        return torch.floor((p - org) / spc)

    pos = wl["pos"].clone()
    # particle j of seed group s is row j * n + s; the work-group of a group is its rank // G
    rank = torch.empty(n, dtype=torch.int64, device="cuda")
    rank[order] = torch.arange(n, device="cuda")
    wg = rank // G
    c0 = cells(pos.view(nb, n, dim).mean(0))
    n_wg = int(wg.max().item()) + 1
    centre = torch.zeros(n_wg, dim, dtype=torch.float64, device="cuda")
    centre.index_add_(0, wg, c0)
    centre = torch.floor(centre / torch.bincount(wg, minlength=n_wg)[:, None].double())
    ctr = centre[wg]  # (n, dim) per seed group
    half = side // 2
    chunk = max(1, wl["steps"] // 20)
    rows = []
    done = 0
    while done < wl["steps"]:
        k = min(chunk, wl["steps"] - done)
        if wl["time_mode"] == "steady":
            W.advect(field, pos, method="rk4", time_mode="steady", n_steps=k, h=wl["h"])
        else:
            W.advect(field, pos, method="rk4", time_mode="linear", n_steps=k, h=wl["h"],
                     t_start=wl["t_start"] + done * wl["h"])
        done += k
        d = cells(pos.view(nb, n, dim)) - ctr[None]
        d = torch.where(per, torch.remainder(d + ncell / 2, ncell) - ncell / 2, d)
        inside = (d.abs() <= half).all(dim=2).float().mean().item()
        rows.append({"step": done, "inside": inside})
    torch.cuda.synchronize()
    mean_inside = sum(r["inside"] for r in rows) / len(rows)
    log(f"tile {side}^{dim} cells: inside fraction over the window " +
        ", ".join(f"{r['step']}:{r['inside']:.3f}" for r in rows))
    print(json.dumps({"workload": wl["name"], "tile_cells_per_axis": side, "record_bytes": rec,
                      "groups_per_block": G, "seed_order": "morton", "chunk_steps": chunk,
                      "inside_fraction_by_step": rows, "mean_inside_fraction": mean_inside}), flush=True)


def run_mlp_train(torch, wl, args):
    """Training steps of one member: forward, backward and Adam per minibatch, timed with
    HIP events over whole epochs; flops = forward + input gradients + weight gradients."""
    S = wl["S"]
    spec, batch = wl["spec"], wl["batch"]
    m = S.FlowMapMLP(spec, seed=1)
    epochs = max(2, args.steps)
    fused = not args.train_layered
    sp = args.train_splits
    S.train(m, wl["X"], wl["D"], epochs=1, batch=batch, lr=3e-3, graph=args.train_graph, fused=fused,
            splits=sp)  # warm-up
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    losses = S.train(m, wl["X"], wl["D"], epochs=epochs, batch=batch, lr=3e-3, graph=args.train_graph, fused=fused,
                     splits=sp)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b)
    steps = epochs * (wl["X"].shape[0] // batch)
    flops = S.mlp_train_flops_per_row(spec) * batch * steps
    tf = flops / (ms * 1e-3) / 1e12
    log(f"{steps} steps of {batch} rows: {ms / steps:.3f} ms/step, {tf:.1f} TFLOP/s, loss {losses[-1]:.3e}")
    print(json.dumps({"metric": "surrogate training TFLOP/s", "value": tf, "unit": "TFLOP/s", "n_gpus": 1,
                      "steps": steps, "ms_per_step": ms / steps, "higher_is_better": True, "dtype": "bf16",
                      "data": "synthetic", "config": dict(wl["config"], graph=bool(args.train_graph),
                                                          kernels="layered" if args.train_layered else "fused",
                                                          splits=sp or "default"),
                      "mfma": {"achieved": tf, "peak": 2500.0, "unit": "TFLOP/s", "frac": tf / 2500.0,
                               "flops_per_row": S.mlp_train_flops_per_row(spec)},
                      "final_loss": losses[-1]}), flush=True)


WORKLOADS = {"abc128": build_abc128, "gyre512": build_gyre, "batch64": build_batch64,
             "mlp_train": build_mlp_train,
             "surrogate2d": build_surrogate2d, "surrogate_infer": build_surrogate_infer,
             "surrogate3d_infer": build_surrogate3d_infer,
             "surrogate3d": build_surrogate3d}


# ---------------------------------------------------------------------------
# accuracy and CPU baseline (oracle used only as checker / baseline)
# ---------------------------------------------------------------------------

def accuracy_vs_fp64(torch, W, wl, ftle32, cfg):
    """Same window in FP64 on the GPU (the kernel that matches the oracle bit
    for bit on small cases), over the full field."""
    from pyftle_amd import _native as N
    from pyftle_amd.field import DeviceField

    f32 = wl["field"]
    # the node layout of the same time levels (the cell layouts hold FP32 records only)
    lay = N.LAYOUT_PAIR if f32.layout in (N.LAYOUT_PAIR, N.LAYOUT_CELL_PAIR) else N.LAYOUT_SINGLE
    f64 = DeviceField(f32.geom, dtype="float64", layout=lay, ring_slots=f32.ring_slots,
                      n_snapshots=f32.n_snapshots, t0=f32.t0, dt_snap=f32.dt_snap)
    if wl["name"] == "abc128":
        f64.pack(0, wl["host"]["vel"])
    else:
        for k, v in enumerate(wl["host"]["snaps"]):
            f64.pack(k, v)
    ftle64 = W.advect_ftle(f64, wl["pos"], method="rk4", time_mode=wl["time_mode"], n_steps=wl["steps"],
                           h=wl["h"], t_start=wl["t_start"], map_period=wl["map_period"], cfg=cfg)
    torch.cuda.synchronize()
    a = ftle32.double()
    b = ftle64
    rng = (b.max() - b.min()).item()
    err = (a - b).abs().max().item()
    del f64
    return {"reference": "fp64 fused kernel (bit-exact with the oracle on the parity cases)",
            "max_abs_err": err, "field_range": rng, "max_abs_err_over_range": err / rng,
            "rmse_over_range": ((a - b).pow(2).mean().sqrt().item()) / rng}


def accuracy_vs_oracle(torch, wl, ftle32, n_sample=32, steps=None):
    """FP64 NumPy oracle on a few seeds (the true checker)."""
    import numpy as np

    from oracle import numpy_oracle as O

    pos = wl["pos"].cpu().numpy()
    n = wl["n_seeds"]
    nb = 2 * wl["dim"]
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(n, size=n_sample, replace=False))
    sub = np.concatenate([pos[j * n + idx] for j in range(nb)], axis=0)
    if wl["name"] == "abc128":
        h = wl["host"]
        nn = h["n"]
        comps = [h["vel"][c].reshape(nn, nn, nn).transpose(2, 1, 0) for c in range(3)]  # (x, y, z)
        gf = O.GridField.__new__(O.GridField)
        gf.dim, gf.v, gf.mins, gf.deltas, gf.periodic = 3, [np.ascontiguousarray(c) for c in comps], \
            [0.0] * 3, [h["sp"]] * 3, (True, True, True)
        exp, _ = O.run_window_linear(lambda t: gf.interpolate, sub, 0.0, wl["h"], wl["steps"], "rk4")
    else:
        h = wl["host"]
        fields = []
        for v in h["snaps"]:
            gf = O.GridField.__new__(O.GridField)
            gf.dim = 2
            gf.v = [np.ascontiguousarray(v[c].reshape(h["ny"], h["nx"]).T) for c in range(2)]
            gf.mins = [0.0, 0.0]
            gf.deltas = [2.0 / (h["nx"] - 1), 1.0 / (h["ny"] - 1)]
            gf.periodic = (False, False)
            fields.append(gf)
        vat = O.time_linear_velocity([f.interpolate for f in fields], 0.0, 0.1)
        exp, _ = O.run_window_linear(vat, sub, 0.0, wl["h"], wl["steps"], "rk4")
    got = ftle32.cpu().numpy()[idx]
    return {"oracle_seeds": int(n_sample), "max_abs_err_vs_oracle": float(np.max(np.abs(got - exp)))}


def cpu_baseline(wl, budget_s=20.0):
    """Reference CPU path (its own compiled GridInterp.cpp under a NumPy
    restatement of the RK4 driver), single thread, on a random seed sample
    sized from a short calibration run so that the whole window takes about
    ``budget_s`` seconds."""
    import numpy as np

    from oracle import native
    from oracle.reference_path import ReferenceGridPath

    if native.reference_ginterp() is None:
        return None
    pos = wl["pos"].cpu().numpy()
    n = wl["n_seeds"]
    nb = 2 * wl["dim"]
    h = wl["host"]
    if wl["dim"] == 3:
        nn = h["n"]
        comps = [h["vel"][c].reshape(nn, nn, nn).transpose(2, 1, 0) for c in range(3)]
        path = ReferenceGridPath(comps, [0.0] * 3, [h["sp"]] * 3)
    else:
        v = h["snaps"][0]
        comps = [v[c].reshape(h["ny"], h["nx"]).T for c in range(2)]
        path = ReferenceGridPath(comps, [0.0, 0.0], [2.0 / (h["nx"] - 1), 1.0 / (h["ny"] - 1)])
    rng = np.random.default_rng(11)

    def sample(k):
        idx = np.sort(rng.choice(n, size=min(k, n), replace=False))
        return idx, np.ascontiguousarray(np.concatenate([pos[j * n + idx] for j in range(nb)], axis=0))

    idx, sub = sample(1024)
    t0 = time.perf_counter()
    path.rk4_steps(sub, wl["h"], 5)
    rate = len(idx) * 5 / (time.perf_counter() - t0)  # seed-steps/s
    k = int(max(256, min(n, budget_s * rate / wl["steps"])))
    idx, sub = sample(k)
    t0 = time.perf_counter()
    path.rk4_steps(sub, wl["h"], wl["steps"])
    el = time.perf_counter() - t0
    seed_steps = len(idx) * wl["steps"] / el
    out = {
        "value": seed_steps / wl["steps"],
        "unit": "seeds/s",
        "cores": 1,
        "kind": "reference",
        "sample": f"{len(idx)} random seeds x the full {wl['steps']}-step {wl['name']} window "
                  f"({el:.1f} s, one thread): the reference GridInterp.cpp compiled -O3 under "
                  f"a NumPy restatement of its RK4 driver (frozen field, zero outside the grid)",
        "seed_steps_per_s": seed_steps,
        "cpu_model": _cpu_model(),
        "host_nproc": os.cpu_count(),
    }
    omp = _cpu_baseline_openmp(wl, comps, path.mins, path.deltas, pos, n, nb, budget_s / 2)
    if omp is not None:
        out["openmp"] = omp
    return out


def _cpu_baseline_openmp(wl, comps, mins, deltas, pos, n, nb, budget_s):
    """The same workload on all host cores: the oracle's C restatement of the reference
    CPU path (cpu_path.c: per-component index conversion and scalar multilinear loop,
    RK4 with full stage arrays), OpenMP over the particles."""
    import numpy as np

    from oracle import native

    try:
        native.cpu_path()
    except RuntimeError:
        return None
    threads = max(1, min(16, os.cpu_count() or 1))  # the GPU box's CPU share is 16 cores

    class _Field:  # the attributes cpu_advect reads
        pass

    f = _Field()
    f.dim, f.v, f.mins, f.deltas = len(comps), [np.ascontiguousarray(c) for c in comps], list(mins), list(deltas)
    rng = np.random.default_rng(12)

    def sample(k):
        idx = np.sort(rng.choice(n, size=min(k, n), replace=False))
        return idx, np.ascontiguousarray(np.concatenate([pos[j * n + idx] for j in range(nb)], axis=0))

    idx, sub = sample(16384)
    t0 = time.perf_counter()
    native.cpu_advect([f], sub, 0.0, wl["h"], 25, "rk4", mode=1, threads=threads)
    rate = len(idx) * 25 / (time.perf_counter() - t0)
    k = int(max(1024, min(n, budget_s * rate / wl["steps"])))
    idx, sub = sample(k)
    t0 = time.perf_counter()
    native.cpu_advect([f], sub, 0.0, wl["h"], wl["steps"], "rk4", mode=1, threads=threads)
    el = time.perf_counter() - t0
    return {
        "value": len(idx) / el,
        "unit": "seeds/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{len(idx)} random seeds x the full {wl['steps']}-step window ({el:.1f} s, "
                  f"{threads} OpenMP threads): the C restatement of the reference CPU path "
                  f"(oracle/cpu_path.c, -O3), same frozen field",
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


# ---------------------------------------------------------------------------

def run_batch(torch, wl, args):
    """64 uneven double-gyre windows in one persistent, graph-captured launch (its bench line)."""
    print(json.dumps(measure_batch(torch, wl, args)), flush=True)


def measure_batch(torch, wl, args, with_cpu=True):
    """64 uneven double-gyre windows in one persistent, graph-captured launch: timed replays,
    the roofline of the binding unit, parity with single-window launches and FP64, and
    (with_cpu) the CPU baseline; returns the bench-line object."""
    from pyftle_amd import window as W
    from pyftle_amd.batch import BatchConfig, BatchPlan
    from pyftle_amd.tuning import best_batch_config

    cfg = best_batch_config("batch64") or BatchConfig()
    if args.no_graph:
        cfg.use_graph = 0
    # seed groups of the large windows in Morton order (set up once with the plan, like its
    # work list; results are identical)
    t_order = time.perf_counter()
    for w in wl["windows"]:
        n = w.positions.shape[0] // 4
        use = args.seed_order == "morton" or (args.seed_order == "auto" and n >= 1 << 18)
        w.order = W.seed_order(w.field, w.positions, 4) if use else None
    torch.cuda.synchronize()
    t_order = time.perf_counter() - t_order
    plan = BatchPlan(wl["windows"], method="rk4", time_mode="linear", cfg=cfg)
    info = plan.info()
    log(f"batch: {info}")
    for _ in range(args.warmup):
        plan.run()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        plan.run()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in evs]
    avg_ms = sum(kernel_ms) / len(kernel_ms)
    log(f"launches (ms): {', '.join(f'{x:.2f}' for x in kernel_ms)}; wall {wall:.3f} s")
    alg = wl["bytes_per_seed_step"] * wl["seed_steps"]
    achieved = alg / (avg_ms * 1e-3) / 1e9
    traffic = load_traffic("batch64")
    hbm_bytes = traffic.get("hbm_bytes_per_launch") if traffic else None
    hbm = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "algorithmic_bytes_per_launch": alg,
           "bytes_per_seed_step": wl["bytes_per_seed_step"], "achieved_algorithmic": achieved,
           "frac_algorithmic": achieved / HBM_PEAK_GBS, "measured_bytes_per_launch": hbm_bytes,
           "measured": hbm_bytes / (avg_ms * 1e-3) / 1e9 if hbm_bytes else None,
           "frac_measured": hbm_bytes / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if hbm_bytes else None}
    kernel_name = "batch_kernel<2,RK4,LINEAR,PAIR> (persistent, hipGraph replay)"
    if plan.windows[0].field.layout in (W.N.LAYOUT_CELL, W.N.LAYOUT_CELL_PAIR):
        # the binding unit, as for the single-window workloads: issued 16-byte record wave
        # loads (counted by a copy of the persistent kernel over the same work list, results
        # identical) x 1 KiB over the timed launch time, against the same path's measured peak
        log("counting issued record loads; probing the vector-memory peak")
        before = [w.ftle.clone() for w in wl["windows"]]
        loads = plan.count_loads()
        identical = all(bool(torch.equal(b, w.ftle)) for b, w in zip(before, wl["windows"]))
        del before
        probe = W.probe_load_rate()
        issued = loads * 1024 / (avg_ms * 1e-3) / 1e9
        peak = probe["bytes_per_s"] / 1e9
        roof = {"bound": "td", "achieved": issued, "peak": peak, "unit": "GB/s", "frac": issued / peak,
                "traffic": hbm_bytes, "kernel": kernel_name, "avg_launch_ms": avg_ms,
                "issued_wave_loads_per_launch": loads, "bytes_per_wave_load": 1024,
                "counting_launch_identical": identical,
                "peak_probe": {"kernel": "load_rate_kernel (coalesced, L1-resident 16-byte wave loads)",
                               "wave_loads": probe["wave_loads"], "ms": probe["seconds"] * 1e3,
                               "wave_loads_per_s": probe["wave_loads_per_s"]},
                "hbm": hbm}
    else:
        roof = dict(hbm, achieved=achieved, frac=achieved / HBM_PEAK_GBS, traffic=hbm_bytes, kernel=kernel_name,
                    avg_launch_ms=avg_ms)
    result = {
        "metric": METRIC, "value": wl["n_seeds"] * args.steps / wall, "unit": "seeds/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic", "config": wl["config"],
        "seed_steps_per_s": wl["seed_steps"] * args.steps / wall,
        "roofline": roof,
        "launch": dict(vars(cfg), **info, seed_order_windows=sum(w.order is not None for w in wl["windows"]),
                       seed_order_setup_s=t_order),
    }
    if not args.no_accuracy:
        log("batch parity against single-window launches and FP64")
        same = True
        worst = 0.0
        checked = 0
        for i, w in enumerate(wl["windows"]):
            single = W.advect_ftle(w.field, w.positions, method="rk4", time_mode="linear",
                                   n_steps=w.n_steps, h=w.h, map_period=w.map_period)
            same = same and bool(torch.equal(single, w.ftle))
            if checked < 6 and w.positions.shape[0] // 4 <= (1 << 17):
                row = wl["table"][i]
                f64 = _gyre_field_on_device(torch, row["grid"][0], row["grid"][1], w.field.n_snapshots,
                                            row["eps"], row["omega"], dtype="float64")
                ref = W.advect_ftle(f64, w.positions, method="rk4", time_mode="linear",
                                    n_steps=w.n_steps, h=w.h, map_period=w.map_period)
                rng_ = (ref.max() - ref.min()).item()
                worst = max(worst, (w.ftle - ref).abs().max().item() / rng_)
                checked += 1
                del f64
        result["accuracy"] = {"batch_equals_single_window_launches": same,
                              "fp64_fields_checked": checked, "max_abs_err_over_range": worst}
    cpu = None
    if with_cpu and not args.no_cpu:
        # the reference CPU path on the batch's most common window shape (512x256 gyre,
        # T = 10), converted to the batch's seeds/s through its mean steps per seed
        cpu = cpu_baseline(_gyre_cpu_workload(torch, {"config": {"grid": [512, 256]}, "h": 0.05, "T": 10.0}),
                           budget_s=15.0)
        if cpu is not None:
            cpu["value"] = cpu["seed_steps_per_s"] * wl["n_seeds"] / wl["seed_steps"]
            cpu["sample"] += (f"; batch seeds/s = seed-steps/s x {wl['n_seeds']} seeds / "
                              f"{wl['seed_steps']} seed-steps of the batch")
    result["cpu_baseline"] = cpu
    plan.close()
    return result


def run_window(torch, wl, args, with_cpu=True):
    """One single-window workload (abc128, gyre512): timed launches, the roofline of the binding
    unit, accuracy against FP64 and the oracle, and (with_cpu) the CPU baseline; returns the
    bench-line object."""
    from pyftle_amd import window as W
    from pyftle_amd.tuning import best_config, shape_class

    cfg = best_config(shape_class(wl["dim"], wl["n_seeds"], wl["time_mode"], records=_records_name(wl["dim"])))
    if args.groups_per_block > 0:
        cfg.groups_per_block = args.groups_per_block
    if args.lane_order >= 0:
        cfg.lane_order = args.lane_order
    if args.xcd_remap >= 0:
        cfg.xcd_remap = args.xcd_remap
    ftle = torch.empty(wl["n_seeds"], dtype=torch.float64, device="cuda")
    nb = 2 * wl["dim"]

    use_order = args.seed_order == "morton" or (args.seed_order == "auto" and wl["n_seeds"] >= 1 << 18)

    def order():
        # part of every step (inside the wall-clock time, outside the kernel's events)
        return W.seed_order(wl["field"], wl["pos"], nb) if use_order else None

    def launch(o):
        W.advect_ftle(wl["field"], wl["pos"], method="rk4", time_mode=wl["time_mode"],
                      n_steps=wl["steps"], h=wl["h"], t_start=wl["t_start"],
                      map_period=wl["map_period"], ftle_out=ftle, cfg=cfg, order=o)

    for _ in range(args.warmup):
        launch(order())
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()  # the stream the kernels are enqueued on
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        o = order()
        a.record(stream)
        launch(o)
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in evs]
    avg_ms = sum(kernel_ms) / len(kernel_ms)
    log(f"launches (ms): {', '.join(f'{x:.2f}' for x in kernel_ms)}; wall {wall:.3f} s")

    n_seeds = wl["n_seeds"]
    value = n_seeds * args.steps / wall
    seed_steps = n_seeds * wl["steps"]
    alg_bytes = wl["bytes_per_seed_step"] * seed_steps
    achieved_alg = alg_bytes / (avg_ms * 1e-3) / 1e9
    traffic = load_traffic(wl["name"])
    hbm_bytes = traffic.get("hbm_bytes_per_launch") if traffic else None
    hbm = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "algorithmic_bytes_per_launch": alg_bytes, "bytes_per_seed_step": wl["bytes_per_seed_step"],
           "achieved_algorithmic": achieved_alg, "frac_algorithmic": achieved_alg / HBM_PEAK_GBS,
           "measured_bytes_per_launch": hbm_bytes,
           "measured": hbm_bytes / (avg_ms * 1e-3) / 1e9 if hbm_bytes else None,
           "frac_measured": hbm_bytes / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if hbm_bytes else None}
    kernel_name = (f"advect_kernel_cell<3,RK4,STEADY,FTLE> ({_records_name(3)} records)" if wl["dim"] == 3
                   else f"advect_kernel_cell<2,RK4,LINEAR,PAIR,FTLE> ({_records_name(2)} records)")
    if _records_name(wl["dim"]) == "cell":
        # The binding unit: the vector-memory (texture) path that serves the record gathers.
        # Issued wave-level 16-byte record loads of one launch, counted by a copy of the kernel
        # (same inputs, launch shape and seed order; results identical), times 1 KiB per wave
        # load, over the timed launch time; peak = the same path's rate for coalesced L1-resident
        # 16-byte wave loads, measured now by pyftle_probe_load_rate.
        log("counting issued record loads; probing the vector-memory peak")
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        ftle_c = torch.empty_like(ftle)
        W.advect_ftle(wl["field"], wl["pos"], method="rk4", time_mode=wl["time_mode"], n_steps=wl["steps"],
                      h=wl["h"], t_start=wl["t_start"], map_period=wl["map_period"], ftle_out=ftle_c, cfg=cfg,
                      order=order(), count_loads=cnt)
        torch.cuda.synchronize()
        loads = int(cnt.item())
        probe = W.probe_load_rate()
        issued = loads * 1024 / (avg_ms * 1e-3) / 1e9
        peak = probe["bytes_per_s"] / 1e9
        roof = {"bound": "td", "achieved": issued, "peak": peak, "unit": "GB/s", "frac": issued / peak,
                "traffic": hbm_bytes, "kernel": kernel_name, "avg_launch_ms": avg_ms,
                "issued_wave_loads_per_launch": loads, "bytes_per_wave_load": 1024,
                "counting_launch_identical": bool(torch.equal(ftle_c, ftle)),
                "peak_probe": {"kernel": "load_rate_kernel (coalesced, L1-resident 16-byte wave loads)",
                               "wave_loads": probe["wave_loads"], "ms": probe["seconds"] * 1e3,
                               "wave_loads_per_s": probe["wave_loads_per_s"]},
                "hbm": hbm}
    else:
        roof = {"bound": "hbm", "achieved": achieved_alg, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved_alg / HBM_PEAK_GBS, "traffic": hbm_bytes, "kernel": kernel_name,
                "algorithmic_bytes_per_launch": alg_bytes, "bytes_per_seed_step": wl["bytes_per_seed_step"],
                "avg_launch_ms": avg_ms}

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "seeds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": wl["config"],
        "seed_steps_per_s": seed_steps * args.steps / wall,
        "roofline": roof,
        "launch": {"groups_per_block": cfg.groups_per_block, "lane_order": cfg.lane_order,
                   "xcd_remap": cfg.xcd_remap, "seed_order": "morton" if use_order else "none"},
    }
    if not args.no_accuracy:
        log("accuracy against FP64")
        acc = accuracy_vs_fp64(torch, W, wl, ftle, cfg)
        acc.update(accuracy_vs_oracle(torch, wl, ftle))
        result["accuracy"] = acc
    result["cpu_baseline"] = None
    if with_cpu and not args.no_cpu:
        log("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(wl)
    return result


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="abc128", choices=sorted(WORKLOADS))
    ap.add_argument("--no-graph", action="store_true",
                    help="batch64: plain launches instead of the captured hipGraph (for per-dispatch counters)")
    ap.add_argument("--surrogate-epochs", type=int, default=0,
                    help="surrogate2d / surrogate3d: training epochs (0 = the workload's default)")
    ap.add_argument("--surrogate-members", type=int, default=0,
                    help="surrogate2d / surrogate3d: ensemble members trained and gated on (0 = the 4 "
                         "of the specification; other counts are a speed/accuracy exploration)")
    ap.add_argument("--surrogate-depth", type=int, default=0,
                    help="surrogate2d / surrogate3d: hidden layers of width 256 / 512 (0 = the 5 / 6 of the "
                         "specified networks; other depths are an accuracy/speed exploration)")
    ap.add_argument("--train-batch", type=int, default=0, help="mlp_train: rows per step (0 = 16384)")
    ap.add_argument("--train-graph", type=int, default=0, help="mlp_train: replay each step as a HIP graph")
    ap.add_argument("--train-shape", default="2d", choices=["2d", "3d"], help="mlp_train: the 2D or the 3D network")
    ap.add_argument("--train-layered", action="store_true", help="mlp_train: the layer-by-layer kernels")
    ap.add_argument("--train-splits", type=int, default=0,
                    help="mlp_train: split-K of the fused step's weight-gradient GEMMs (0 = its default)")
    ap.add_argument("--tile-occupancy", action="store_true",
                    help="abc128 / gyre512: only measure how much of the window an LDS tile per work-group "
                         "could serve (share of particles inside their work-group's tile, by step)")
    ap.add_argument("--issue-probe", action="store_true",
                    help="only measure the issue cost (cycles) of MFMA, FP32 multiply, exp2 and reciprocal, "
                         "alone and with an MFMA wave and a VALU wave sharing a SIMD")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-other-workloads", action="store_true",
                    help="default line: skip its gyre512 and batch64 objects")
    ap.add_argument("--no-surrogate3d", action="store_true",
                    help="default line: skip its 3D surrogate leg (the 2D leg stays)")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="abc128: skip the surrogate leg (training and timing the 2D surrogate, ~1 min)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 accuracy check")
    ap.add_argument("--seed-order", default="auto", choices=["auto", "morton", "none"],
                    help="abc128 / gyre512: process seed groups in Morton order of their cells (results are "
                         "identical; computed inside every timed step); auto = from 2^18 seed groups")
    ap.add_argument("--records", default="auto", choices=["auto", "node", "cell"],
                    help="record layout of the FP32 fields of abc128, gyre512, batch64 and the 2D surrogate "
                         "workloads: one record per grid node, one per cell holding its corners, or auto "
                         "(cell records)")
    ap.add_argument("--groups-per-block", type=int, default=0)
    ap.add_argument("--lane-order", type=int, default=-1)
    ap.add_argument("--xcd-remap", type=int, default=-1)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only copy this many bytes device-to-device 3 times (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every FTLE seed group is independent and this build "
              "runs on a single MI355X; --gpus above 1 is not supported.", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4}), flush=True)
        return

    from pyftle_amd import _native
    from pyftle_amd import window as W
    from pyftle_amd.tuning import best_config, shape_class

    _native.load()
    torch.cuda.set_device(0)
    if args.issue_probe:
        print(json.dumps({"issue_probe": W.probe_issue_rates()}), flush=True)
        return
    log(f"building workload {args.workload}")
    kw = {}
    if args.workload in ("surrogate2d", "surrogate3d") and args.surrogate_epochs > 0:
        kw["epochs"] = args.surrogate_epochs
    if args.workload in ("surrogate2d", "surrogate3d") and args.surrogate_members > 0:
        kw["members"] = args.surrogate_members
    if args.workload in ("surrogate2d", "surrogate3d") and args.surrogate_depth > 0:
        kw["depth"] = args.surrogate_depth
    if args.workload == "mlp_train":
        kw["shape"] = args.train_shape
        if args.train_batch > 0:
            kw["batch"] = args.train_batch
    global RECORDS
    RECORDS = args.records
    wl = WORKLOADS[args.workload](torch, **kw)
    if "config" in wl and args.workload not in ("surrogate3d", "surrogate3d_infer", "mlp_train"):
        wl["config"]["records"] = _records_name(wl.get("dim", 2))
    if wl["name"] == "batch64":
        return run_batch(torch, wl, args)
    if wl["name"] == "mlp_train":
        return run_mlp_train(torch, wl, args)
    if args.tile_occupancy:
        return run_tile_occupancy(torch, wl, args)
    if wl["name"] == "surrogate2d":
        return run_surrogate(torch, wl, args)
    if wl["name"] == "surrogate_infer":
        return run_surrogate_infer(torch, wl, args)
    if wl["name"] == "surrogate3d_infer":
        return run_surrogate3d_infer(torch, wl, args)
    if wl["name"] == "surrogate3d":
        return run_surrogate3d(torch, wl, args)
    result = run_window(torch, wl, args)
    if wl["name"] == "abc128" and not args.no_surrogate:
        # the metric's surrogate half: the 2D flow-map surrogate against plain GPU RK4
        log("surrogate leg (2D double gyre, 4 members of 5x256, 67.1M seed-windows)")
        del wl
        torch.cuda.empty_cache()
        summ, plan = surrogate2d_measure(torch, build_surrogate2d(torch), args)
        plan.close()
        result["surrogate"] = summ
        del summ, plan
        torch.cuda.empty_cache()
        if not args.no_surrogate3d:
            # and the 3D surrogate (unsteady ABC 192^3, 4 members of 6x512) on one 256^3-seed window
            log("surrogate leg (3D unsteady ABC, 4 members of 6x512, 16.8M seeds)")
            summ3, _ = surrogate3d_measure(torch, build_surrogate3d(torch), args, full=False)
            result["surrogate3d"] = summ3
            del summ3
            torch.cuda.empty_cache()
    if args.workload == "abc128" and not args.no_other_workloads:
        # the other plain-RK4 configurations, so that every figure of the headline is on this line:
        # the 2D double-gyre window (C0) and the 64-field batch (C2), without their CPU baselines
        keep = ("config", "ms_per_step", "value", "unit", "seed_steps_per_s", "roofline", "launch", "accuracy")
        for name, build, measure in (("gyre512", build_gyre, run_window), ("batch64", build_batch64, measure_batch)):
            log(f"workload {name}")
            w2 = build(torch)
            w2["config"]["records"] = _records_name(2)
            r = measure(torch, w2, args, with_cpu=False)
            result[name] = {k: r[k] for k in keep if k in r}
            del w2, r
            torch.cuda.empty_cache()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
