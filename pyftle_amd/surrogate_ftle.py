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
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass

import torch

from . import window as W
from .batch import BatchConfig, BatchPlan, Window
from .field import DeviceField
from .surrogate import FlowMapEnsemble, FlowMapMLP, MLPSpec, ftle_from_displacement_jacobian, train


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
                 lr: float = 2e-3, log=None) -> FlowMapEnsemble:
    models = []
    for k in range(members):
        m = FlowMapMLP(spec, seed=1000 + k)
        t0 = time.perf_counter()
        losses = train(m, X, D, epochs=epochs, batch=batch, lr=lr, seed=k)
        torch.cuda.synchronize()
        if log:
            log(f"member {k}: loss {losses[0]:.3e} -> {losses[-1]:.3e} in {time.perf_counter() - t0:.1f} s")
        models.append(m)
    return FlowMapEnsemble(models)


@dataclass
class SurrogateResult:
    ftle: torch.Tensor          # (N,) FP64
    fallback: torch.Tensor      # (N,) bool: seeds recomputed by RK4
    spread: torch.Tensor        # (N,) member disagreement


class SurrogateFTLE:
    """FTLE of seeds at window start t0 from the ensemble, RK4 where it disagrees."""

    def __init__(self, ensemble: FlowMapEnsemble, field: DeviceField, T: float, h: float,
                 delta: float = 1e-3, tol: float = 1e-3, time_mode: str = "linear", chunk: int = 1 << 21,
                 jacobian_members: int | None = None, gate: str = "ftle"):
        if gate not in ("ftle", "displacement"):
            raise ValueError(f"unknown gate {gate!r}: 'ftle' or 'displacement'")
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

    def predict(self, seeds: torch.Tensor, t0: float):
        """Surrogate-only FTLE and the members' disagreement (``gate``) for ``seeds`` (N, dim)."""
        dim = seeds.shape[1]
        out_f, out_s = [], []
        for s0 in range(0, seeds.shape[0], self.chunk):
            sd = seeds[s0:s0 + self.chunk]
            x = torch.empty(sd.shape[0], dim + 1, dtype=torch.float32, device=sd.device)
            x[:, :dim] = sd.float()
            x[:, dim] = t0
            if self.gate == "ftle":
                _, Jm, _ = self.ens.predict(x, dim, per_member=True)
                f = ftle_from_displacement_jacobian(Jm.mean(0), self.T)
                dev = torch.zeros_like(f)
                for k in range(Jm.shape[0]):
                    torch.maximum(dev, (ftle_from_displacement_jacobian(Jm[k], self.T) - f).abs(), out=dev)
                out_f.append(f)
                out_s.append(dev)
            else:
                _, Jd, spread = self.ens.predict(x, dim, jacobian_members=self.jacobian_members)
                out_f.append(ftle_from_displacement_jacobian(Jd, self.T))
                out_s.append(spread)
        f = torch.cat(out_f)
        spread = torch.cat(out_s)
        if self.gate == "ftle":  # in units of the window's FTLE range
            fin = f[torch.isfinite(f)]
            rng = (fin.max() - fin.min()).clamp_min(1e-300) if fin.numel() else torch.ones((), device=f.device)
            spread = spread / rng
            spread[~torch.isfinite(spread)] = float("inf")
        return f, spread

    def rk4_ftle(self, seeds: torch.Tensor, t0: float) -> torch.Tensor:
        """Plain RK4 FTLE of the same seeds (neighbour groups at +-delta), with the launch
        shape tuned for this field's shape class when the tuning cache holds one."""
        pos = neighbour_groups(seeds, self.delta)
        steps = int(round(self.T / abs(self.h)))
        return W.advect_ftle(self.field, pos, method="rk4", time_mode=self.time_mode, n_steps=steps,
                             h=self.h, t_start=t0, map_period=steps * abs(self.h),
                             order="morton" if seeds.shape[0] >= 1 << 18 else None)

    def ftle(self, seeds: torch.Tensor, t0: float) -> SurrogateResult:
        f, spread = self.predict(seeds, t0)
        bad = spread > self.tol
        idx = torch.nonzero(bad).flatten()
        if idx.numel():
            f[idx] = self.rk4_ftle(seeds[idx].contiguous(), t0)
        return SurrogateResult(ftle=f, fallback=bad, spread=spread)

    def ftle_windows(self, seeds: torch.Tensor, t0s, cfg: BatchConfig | None = None) -> list[SurrogateResult]:
        """``ftle`` for several window starts: the ensemble per start, then the fallback seeds
        of every start in one batched persistent RK4 launch (one window per start) instead of
        one small launch each.  The batched kernel is bit-equal to single launches."""
        steps = int(round(self.T / abs(self.h)))
        results, wins, where = [], [], []
        for t0 in t0s:
            f, spread = self.predict(seeds, t0)
            bad = spread > self.tol
            results.append(SurrogateResult(ftle=f, fallback=bad, spread=spread))
            idx = torch.nonzero(bad).flatten()
            if idx.numel():
                pos = neighbour_groups(seeds[idx].contiguous(), self.delta)
                wins.append(Window(field=self.field, positions=pos, n_steps=steps, h=self.h, t_start=float(t0),
                                   map_period=steps * abs(self.h)))
                where.append((len(results) - 1, idx))
        if wins:
            plan = BatchPlan(wins, method="rk4", time_mode=self.time_mode, cfg=cfg)
            plan.run()
            for w, (k, idx) in zip(wins, where):
                results[k].ftle[idx] = w.ftle
            plan.close()  # waits for the device
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
