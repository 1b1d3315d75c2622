"""FTLE fields from the learned flow-map surrogate, with RK4 as fallback and accuracy reference.

Pipeline:

1. ``make_training_data``: trajectories of the fused RK4 kernel
   (``pyftle_advect``) from random seeds and window starts ``t0``; targets are
   the displacements over the window ``T``.
2. ``fit_ensemble``: several ``FlowMapMLP`` members trained by the HIP
   forward/backward/Adam kernels.
3. ``SurrogateFTLE.ftle``: per seed the members' displacement and its
   Jacobian (JVP tangents) -> FTLE of ``J = I + mean dD/dx`` with the FP64
   closed-form kernels; seeds where the members disagree by more than ``tol`` are
   redone by the fused RK4 window kernel (the fallback), so the returned field is
   RK4 wherever the ensemble is unsure.

The disagreement (``gate``):
* ``"ftle"`` (default): every member's own FTLE (from its own JVP Jacobian) against the
  FTLE of the ensemble's mean Jacobian, ``max_m |ftle_m - ftle|``, in units of the
  window's FTLE range (``max - min`` of the ensemble field over the queried seeds) --
  the quantity the accuracy bar is stated in;
* ``"displacement"``: the largest deviation of a member's displacement from the mean,
  in domain units.

The probe plan (``probe > 0``, large seed sets): the ensemble first runs on every
``probe``-th seed (a spread-out subsample).  The seeds' bounding box is cut into
``tiles`` per axis; a tile where more than ``tile_fallback_max`` of the probed seeds fail
the gate goes to RK4 whole, without the ensemble pass, and the other tiles are evaluated
seed by seed as above.  The FTLE gate's range then comes from the probe.  Every seed
returned from the ensemble has passed its own gate, so the plan only decides where the
ensemble pass is not worth its cost: where RK4 would redo most of a tile anyway (at a
tolerance that fails everywhere, the whole field is RK4 plus the probe's cost).
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass

import torch

from . import window as W
from .batch import BatchConfig, BatchPlan, Window
from .field import DeviceField
from .surrogate import FlowMapEnsemble, FlowMapMLP, MLPSpec, ftle_from_displacement_jacobian, train, train_ensemble


def uniform_seeds(n: int, lower, upper, gen: torch.Generator, device="cuda") -> torch.Tensor:
    lo = torch.tensor(lower, dtype=torch.float64, device=device)
    hi = torch.tensor(upper, dtype=torch.float64, device=device)
    return lo + (hi - lo) * torch.rand(n, len(lower), generator=gen, dtype=torch.float64, device=device)


def make_training_data(field: DeviceField, n_traj: int, T: float, h: float, t0_range, lower, upper,
                       n_bins: int = 512, seed: int = 0, time_mode: str = "linear"):
    """``n_traj`` RK4 trajectories: inputs (x, t0) FP32 and displacements over ``T``.

    Window starts are spread over ``n_bins`` evenly spaced values of
    ``t0_range`` (one advection launch per value).
    """
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev).manual_seed(seed)
    dim = len(lower)
    per = n_traj // n_bins
    steps = int(round(T / abs(h)))
    X = torch.empty(per * n_bins, dim + 1, dtype=torch.float32, device=dev)
    D = torch.empty(per * n_bins, dim, dtype=torch.float32, device=dev)
    for b in range(n_bins):
        t0 = t0_range[0] + (t0_range[1] - t0_range[0]) * (b + 0.5) / n_bins
        x0 = uniform_seeds(per, lower, upper, gen)
        x = x0.clone()
        W.advect(field, x, method="rk4", time_mode=time_mode, n_steps=steps, h=h, t_start=t0)
        X[b * per:(b + 1) * per, :dim] = x0.float()
        X[b * per:(b + 1) * per, dim] = t0
        D[b * per:(b + 1) * per] = (x - x0).float()
    return X, D


def fit_ensemble(X, D, spec: MLPSpec, members: int = 4, epochs: int = 20, batch: int = 32768,
                 lr: float = 2e-3, log=None, concurrent: bool = True) -> FlowMapEnsemble:
    """Members ``FlowMapMLP(spec, seed=1000 + k)`` trained on (X, D), member k shuffling with
    seed k.  ``concurrent`` (default): the members' runs interleaved on a stream each
    (``train_ensemble``); otherwise one after another (``train``).  The weights are the same
    bit for bit either way."""
    models = [FlowMapMLP(spec, seed=1000 + k) for k in range(members)]
    t0 = time.perf_counter()
    if concurrent:
        all_losses = train_ensemble(models, X, D, epochs=epochs, batch=batch, lr=lr, seeds=range(members))
    else:
        all_losses = [train(m, X, D, epochs=epochs, batch=batch, lr=lr, seed=k) for k, m in enumerate(models)]
    torch.cuda.synchronize()
    if log:
        for k, losses in enumerate(all_losses):
            log(f"member {k}: loss {losses[0]:.3e} -> {losses[-1]:.3e}")
        log(f"{members} members trained in {time.perf_counter() - t0:.1f} s"
            f"{' (concurrently)' if concurrent else ''}")
    return FlowMapEnsemble(models)


@dataclass
class SurrogateResult:
    ftle: torch.Tensor          # (N,) FP64
    fallback: torch.Tensor      # (N,) bool: seeds recomputed by RK4
    spread: torch.Tensor        # (N,) member disagreement (NaN where a probe-planned tile skipped the ensemble)


@dataclass
class ProbePlan:
    tile: torch.Tensor          # (N,) tile index of every seed
    probe: torch.Tensor         # (M,) indices of the probed seeds
    n_tiles: int
    tile_seeds: torch.Tensor    # (n_tiles,) seeds per tile, host


# multiplier of the probe's index permutation (odd, and coprime to any seed count it is used
# with: the probe falls back to a plain stride when it is not)
_PROBE_MULT = 2654435761


class SurrogateFTLE:
    """FTLE of seeds at window start t0 from the ensemble, RK4 where it disagrees."""

    def __init__(self, ensemble: FlowMapEnsemble, field: DeviceField, T: float, h: float,
                 delta: float = 1e-3, tol: float = 1e-3, time_mode: str = "linear", chunk: int = 1 << 21,
                 jacobian_members: int | None = None, gate: str = "ftle", probe: int = 256, tiles: int = 16,
                 tile_fallback_max: float = 0.5, plan_min_seeds: int = 1 << 16):
        if gate not in ("ftle", "displacement"):
            raise ValueError(f"unknown gate {gate!r}: 'ftle' or 'displacement'")
        if probe < 0 or tiles < 1 or not 0.0 <= tile_fallback_max <= 1.0:
            raise ValueError("probe >= 0, tiles >= 1 and 0 <= tile_fallback_max <= 1 are required")
        self.probe = probe
        self.tiles = tiles
        self.tile_fallback_max = tile_fallback_max
        self.plan_min_seeds = plan_min_seeds
        self.ens = ensemble
        self.gate = gate
        self.jacobian_members = jacobian_members
        self.field = field
        self.T = T
        self.h = h
        self.delta = delta
        self.tol = tol
        self.time_mode = time_mode
        self.chunk = chunk

    def _evaluate_rows(self, x: torch.Tensor, dim: int):
        """The ensemble's FTLE and raw disagreement for query rows ``x`` = (x, t0), (M, dim + 1)."""
        if self.gate == "ftle":
            _, Jm, _ = self.ens.predict(x, dim, per_member=True)
            f = ftle_from_displacement_jacobian(Jm.mean(0), self.T)
            dev = torch.zeros_like(f)
            for k in range(Jm.shape[0]):
                torch.maximum(dev, (ftle_from_displacement_jacobian(Jm[k], self.T) - f).abs(), out=dev)
            return f, dev
        _, Jd, spread = self.ens.predict(x, dim, jacobian_members=self.jacobian_members)
        return ftle_from_displacement_jacobian(Jd, self.T), spread

    def _evaluate(self, seeds: torch.Tensor, t0s):
        """The ensemble's FTLE and the raw disagreement (FTLE gate: absolute FTLE units;
        displacement gate: domain units) for ``seeds`` (N, dim) at one window start, or at
        each of several (rows ordered start by start), in chunks of ``chunk`` rows."""
        n, dim = seeds.shape
        t0s = list(t0s) if isinstance(t0s, (list, tuple)) else [t0s]
        out_f, out_s = [], []
        for r0 in range(0, n * len(t0s), self.chunk):
            r1 = min(r0 + self.chunk, n * len(t0s))
            x = torch.empty(r1 - r0, dim + 1, dtype=torch.float32, device=seeds.device)
            r = r0
            while r < r1:  # the rows of one window start at a time
                w, s0 = divmod(r, n)
                s1 = min(n, s0 + r1 - r)
                x[r - r0:r - r0 + s1 - s0, :dim] = seeds[s0:s1]
                x[r - r0:r - r0 + s1 - s0, dim] = t0s[w]
                r += s1 - s0
            f, dev = self._evaluate_rows(x, dim)
            out_f.append(f)
            out_s.append(dev)
        return torch.cat(out_f), torch.cat(out_s)

    def _normalise(self, dev: torch.Tensor, f_range: torch.Tensor | None) -> torch.Tensor:
        """The gate's disagreement: the FTLE gate divides by the window's FTLE range."""
        if self.gate != "ftle":
            return dev
        spread = dev / f_range
        spread[~torch.isfinite(spread)] = float("inf")
        return spread

    @staticmethod
    def _range(f: torch.Tensor) -> torch.Tensor:
        """``max - min`` of the finite values (1 when there are none), without a host sync."""
        fin = torch.isfinite(f)
        r = torch.where(fin, f, float("-inf")).max() - torch.where(fin, f, float("inf")).min()
        return torch.where(torch.isfinite(r), r.clamp_min(1e-300), torch.ones_like(r))

    def predict(self, seeds: torch.Tensor, t0: float):
        """Surrogate-only FTLE and the members' disagreement (``gate``) for ``seeds`` (N, dim)."""
        f, dev = self._evaluate(seeds, t0)
        return f, self._normalise(dev, self._range(f) if self.gate == "ftle" else None)

    def _planned(self, seeds: torch.Tensor) -> bool:
        return self.probe > 0 and seeds.shape[0] >= self.plan_min_seeds

    def _tiles_of(self, seeds: torch.Tensor) -> ProbePlan:
        """The probe plan of a seed set: the tile of every seed over the seeds' bounding box
        and the probed seeds (every ``probe``-th seed of a multiplicative permutation).  The
        tiles per axis are ``tiles``, fewer when a tile would hold under 32 probed seeds."""
        n, dim = seeds.shape
        m = max(1, n // self.probe)
        per_axis = max(1, min(self.tiles, int((m / 32.0) ** (1.0 / dim))))
        # the seeds' box in one pass; a non-finite coordinate shows in it (NaN or +-inf), and
        # only then the box is taken over the finite ones (non-finite coordinates fall in tile 0
        # of their axis)
        lo, hi = W.column_box(seeds)
        finite = bool(torch.isfinite(lo).all() and torch.isfinite(hi).all())
        if not finite:
            fin = torch.isfinite(seeds)
            lo = torch.where(fin, seeds, float("inf")).min(0).values
            hi = torch.where(fin, seeds, float("-inf")).max(0).values
            lo = torch.where(torch.isfinite(lo), lo, torch.zeros_like(lo))
        ext = (hi - lo).nan_to_num(1.0, 1.0, 1.0).clamp_min(1e-300)
        cell = torch.sub(seeds, lo).div_(ext).mul_(per_axis)  # one temporary, updated in place
        if not finite:
            cell.nan_to_num_(0.0, 0.0, 0.0)
        cell.floor_().clamp_(0, per_axis - 1)
        strides = torch.tensor([float(per_axis ** (dim - 1 - a)) for a in range(dim)], dtype=cell.dtype,
                               device=seeds.device)
        tid = cell.mul_(strides).sum(1).long()  # exact: small integers
        del cell
        k = torch.arange(m, dtype=torch.long, device=seeds.device)
        idx = (k * _PROBE_MULT) % n if math.gcd(_PROBE_MULT, n) == 1 else k * self.probe
        n_tiles = per_axis ** dim
        return ProbePlan(tile=tid, probe=idx, n_tiles=n_tiles,
                         tile_seeds=torch.bincount(tid, minlength=n_tiles).cpu())

    def _probe(self, seeds: torch.Tensor, t0s, plan):
        """The ensemble on the probed seeds at every window start, in one pass.  Returns, per
        start, the tiles sent to RK4 whole (device and host copies; one host sync for all
        starts) and the FTLE range the gate divides by."""
        tid, probe_idx, n_tiles = plan.tile, plan.probe, plan.n_tiles
        ps = seeds[probe_idx].contiguous()
        m = ps.shape[0]
        n_w = len(t0s)
        f_p, dev_p = self._evaluate(ps, t0s)
        bad_p, ranges = [], []
        for w in range(n_w):
            f_w, d_w = f_p[w * m:(w + 1) * m], dev_p[w * m:(w + 1) * m]
            f_range = self._range(f_w) if self.gate == "ftle" else None
            bad_p.append(self._normalise(d_w, f_range) > self.tol)
            ranges.append(f_range)
        bad_p = torch.stack(bad_p)                               # (starts, probes)
        t_p = tid[probe_idx]
        key = (torch.arange(n_w, device=seeds.device)[:, None] * n_tiles + t_p[None, :])
        cnt = torch.bincount(t_p, minlength=n_tiles).double()    # probes per tile
        fail = torch.bincount(key[bad_p], minlength=n_w * n_tiles).double().view(n_w, n_tiles)
        frac_all = fail.sum(1, keepdim=True) / cnt.sum().clamp_min(1.0)
        frac = torch.where(cnt > 0, fail / cnt.clamp_min(1.0), frac_all)
        tile_rk4 = frac > self.tile_fallback_max                 # (starts, tiles)
        host = tile_rk4.cpu()
        return [(tile_rk4[w], host[w], ranges[w]) for w in range(n_w)]

    def _window(self, seeds: torch.Tensor, t0: float, plan=None, probe=None):
        """One window start: the ensemble's FTLE and the fallback mask before RK4, enqueued
        without a host sync.  With the probe plan (``_tiles_of`` and ``_probe``), tiles the
        probe rates as mostly failing skip the ensemble pass."""
        if plan is None:
            f, spread = self.predict(seeds, t0)
            return f, spread > self.tol, spread
        tid, n_tiles = plan.tile, plan.n_tiles
        tile_rk4, tile_rk4_host, f_range = probe if probe is not None else self._probe(seeds, [t0], plan)[0]
        n_rk4_tiles = int(tile_rk4_host.sum())
        if n_rk4_tiles == 0:  # every seed through the ensemble, no gather
            f, dev = self._evaluate(seeds, t0)
            spread = self._normalise(dev, f_range).to(torch.float64)
            return f, spread > self.tol, spread
        to_rk4 = tile_rk4[tid]
        n = seeds.shape[0]
        f = torch.full((n,), float("nan"), dtype=torch.float64, device=seeds.device)
        spread = torch.full((n,), float("nan"), dtype=torch.float64, device=seeds.device)
        bad = to_rk4.clone()
        if n_rk4_tiles < n_tiles:
            n_kept = int(plan.tile_seeds[~tile_rk4_host].sum())
            idx = torch.nonzero_static(~to_rk4, size=n_kept).flatten()
            f_s, dev_s = self._evaluate(seeds[idx].contiguous(), t0)
            sp = self._normalise(dev_s, f_range).to(torch.float64)
            f[idx] = f_s
            spread[idx] = sp
            bad[idx] = sp > self.tol
        return f, bad, spread

    def rk4_ftle(self, seeds: torch.Tensor, t0: float) -> torch.Tensor:
        """Plain RK4 FTLE of the same seeds (neighbour groups at +-delta), with the launch
        shape tuned for this field's shape class when the tuning cache holds one."""
        pos = neighbour_groups(seeds, self.delta)
        steps = int(round(self.T / abs(self.h)))
        return W.advect_ftle(self.field, pos, method="rk4", time_mode=self.time_mode, n_steps=steps,
                             h=self.h, t_start=t0, map_period=steps * abs(self.h),
                             order="morton" if seeds.shape[0] >= 1 << 18 else None)

    def ftle(self, seeds: torch.Tensor, t0: float) -> SurrogateResult:
        f, bad, spread = self._window(seeds, t0, self._tiles_of(seeds) if self._planned(seeds) else None)
        c = int(bad.sum())
        if c == seeds.shape[0]:  # the whole field from RK4, no gather
            f = self.rk4_ftle(seeds, t0)
        elif c:
            idx = torch.nonzero(bad).flatten()
            f[idx] = self.rk4_ftle(seeds[idx].contiguous(), t0)
        return SurrogateResult(ftle=f, fallback=bad, spread=spread)

    def ftle_windows(self, seeds: torch.Tensor, t0s, cfg: BatchConfig | None = None) -> list[SurrogateResult]:
        """``ftle`` for several window starts: the ensemble per start (enqueued back to back,
        no host sync between starts), then the fallback seeds of every start in one batched
        persistent RK4 launch (one window per start) instead of one small launch each.  The
        batched kernel is bit-equal to single launches."""
        steps = int(round(self.T / abs(self.h)))
        n, dim = seeds.shape
        t0s = [float(t) for t in t0s]
        plan = self._tiles_of(seeds) if self._planned(seeds) else None
        probes = self._probe(seeds, t0s, plan) if plan is not None else [None] * len(t0s)
        results = []
        for t0, probe in zip(t0s, probes):
            f, bad, spread = self._window(seeds, t0, plan, probe)
            results.append(SurrogateResult(ftle=f, fallback=bad, spread=spread))
        if not results:
            return results
        bads = torch.stack([r.fallback for r in results])        # (starts, seeds)
        counts = bads.sum(1).tolist()                            # host sync 1
        # the fallback seeds' indices, needed only where a start falls back in part
        flat = torch.nonzero(bads.view(-1)).flatten() if any(0 < c < n for c in counts) else None  # sync 2
        wins, where = [], []
        all_pos = all_order = None  # the whole seed set's neighbour groups, built once when needed
        off = 0
        for k, (t0, c) in enumerate(zip(t0s, counts)):
            if not c:
                continue
            idx = None  # the whole start
            if flat is not None:
                idx = flat[off:off + c] - k * n
                off += c
            if c == n:
                if all_pos is None:
                    all_pos = neighbour_groups(seeds, self.delta)
                    all_order = W.seed_order(self.field, all_pos, 2 * dim) if n >= 1 << 18 else None
                pos, order = all_pos, all_order
            else:
                pos = neighbour_groups(seeds[idx].contiguous(), self.delta)
                order = W.seed_order(self.field, pos, 2 * dim) if c >= 1 << 18 else None
            wins.append(Window(field=self.field, positions=pos, n_steps=steps, h=self.h, t_start=t0,
                               map_period=steps * abs(self.h), order=order))
            where.append((k, idx))
        if wins:
            batch = BatchPlan(wins, method="rk4", time_mode=self.time_mode, cfg=cfg)
            batch.run()
            for w, (k, idx) in zip(wins, where):
                if idx is None:
                    results[k].ftle.copy_(w.ftle)
                else:
                    results[k].ftle[idx] = w.ftle
            batch.close()  # waits for the device
        return results


def neighbour_groups(seeds: torch.Tensor, delta: float) -> torch.Tensor:
    """Block-ordered [L, R, T, B(, F, Bk)] neighbours (particles.py:170,191,212)."""
    dim = seeds.shape[1]
    blocks = []
    for a in range(dim):
        e = torch.zeros(dim, dtype=seeds.dtype, device=seeds.device)
        e[a] = delta
        blocks += [seeds - e, seeds + e] if a == 0 else [seeds + e, seeds - e]
    return torch.cat(blocks, 0).contiguous()


def rk4_ftle_windows(field: DeviceField, seeds: torch.Tensor, t0s, T: float, h: float, delta: float = 1e-3,
                     time_mode: str = "linear", cfg: BatchConfig | None = None):
    """Plain RK4 FTLE of the same seeds for several window starts in one batched launch."""
    steps = int(round(T / abs(h)))
    pos = neighbour_groups(seeds, delta)
    dim = seeds.shape[1]
    # one Morton order of the shared seeds for every window (results do not depend on it)
    order = W.seed_order(field, pos, 2 * dim) if seeds.shape[0] >= 1 << 18 else None
    wins = [Window(field=field, positions=pos, n_steps=steps, h=h, t_start=float(t0), map_period=steps * abs(h),
                   order=order) for t0 in t0s]
    plan = BatchPlan(wins, method="rk4", time_mode=time_mode, cfg=cfg)
    return plan, wins


def relative_rmse(a: torch.Tensor, b: torch.Tensor) -> float:
    rng = (b.max() - b.min()).item()
    return math.sqrt(((a - b) ** 2).mean().item()) / rng
