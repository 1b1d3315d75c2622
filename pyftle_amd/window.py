"""Window engine: one launch from seed groups to an FTLE field.

``advect_ftle`` wraps ``pyftle_advect_ftle`` (fused advection + FP64 FTLE
epilogue); ``advect`` wraps ``pyftle_advect`` (positions only).  Both take
device tensors and enqueue on the current stream.

Time semantics (``time_mode``):
* ``"frozen"``: step ``i`` reads snapshot ``first + i`` for every stage
  (the reference, ftle_solver.py:99-102 and integrate.py:323-339);
* ``"steady"``: every stage reads snapshot ``first``;
* ``"linear"``: stage time ``t`` blends snapshots ``k`` and ``k+1``
  (needs a ``PAIR``-layout field).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _native as N
from .field import DeviceField, dtype_code, ptr, stream_handle

TIME_MODES = {"steady": N.TIME_STEADY, "frozen": N.TIME_FROZEN, "linear": N.TIME_LINEAR}
METHODS = {"euler": N.EULER, "ab2": N.AB2, "rk4": N.RK4}


@dataclass
class LaunchConfig:
    """Tunable launch shape of the fused kernel (see ``autotune``)."""

    groups_per_block: int = 64
    lane_order: int = N.ORDER_SEED_MAJOR
    xcd_remap: int = 1
    compute: int = N.COMPUTE_NATIVE


def make_opts(method, time_mode, n_steps, h, first=0, t_start=0.0, have_prev=False,
              cfg: LaunchConfig | None = None) -> N.AdvectOpts:
    cfg = cfg or LaunchConfig()
    o = N.AdvectOpts()
    o.method = METHODS[method] if isinstance(method, str) else int(method)
    o.time_mode = TIME_MODES[time_mode] if isinstance(time_mode, str) else int(time_mode)
    o.n_steps = int(n_steps)
    o.first_snapshot = int(first)
    o.t_start = float(t_start)
    o.h = float(h)
    o.ab2_have_prev = int(bool(have_prev))
    o.lane_order = int(cfg.lane_order)
    o.groups_per_block = int(cfg.groups_per_block)
    o.xcd_remap = int(cfg.xcd_remap)
    o.compute = int(cfg.compute)
    return o


def advect_ftle(field: DeviceField, positions: torch.Tensor, *, method="rk4", time_mode="frozen",
                n_steps: int, h: float, map_period: float, first: int = 0, t_start: float = 0.0,
                positions_out: torch.Tensor | None = None, ftle_out: torch.Tensor | None = None,
                cfg: LaunchConfig | None = None, stream=None) -> torch.Tensor:
    """Advect block-ordered seed groups and return the FP64 FTLE per seed."""
    if not positions.is_cuda:
        raise ValueError("positions must be a device tensor")
    positions = positions.contiguous()
    dim = field.geom.dim
    if positions.dim() != 2 or positions.shape[1] != dim:
        raise ValueError(f"positions must have shape (nb*N, {dim})")
    nb = 2 * dim
    if positions.shape[0] % nb != 0:
        raise ValueError(f"positions.shape[0] must be a multiple of {nb}")
    n_groups = positions.shape[0] // nb
    if ftle_out is None:
        ftle_out = torch.empty(n_groups, dtype=torch.float64, device=positions.device)
    opts = make_opts(method, time_mode, n_steps, h, first, t_start, False, cfg)
    pout = ptr(positions_out) if positions_out is not None else ctypes.c_void_p(0)
    N.call("pyftle_advect_ftle", ctypes.byref(field.desc), ctypes.byref(opts), ptr(positions),
           dtype_code(positions.dtype), n_groups, pout, float(map_period), ptr(ftle_out),
           stream_handle(stream))
    return ftle_out


def advect(field: DeviceField, positions: torch.Tensor, *, method="rk4", time_mode="frozen",
           n_steps: int, h: float, first: int = 0, t_start: float = 0.0, vprev=None,
           have_prev=False, cfg: LaunchConfig | None = None, stream=None) -> torch.Tensor:
    """Advect particles ``(P, dim)`` in place for ``n_steps``."""
    if not positions.is_cuda or not positions.is_contiguous():
        raise ValueError("positions must be a contiguous device tensor")
    opts = make_opts(method, time_mode, n_steps, h, first, t_start, have_prev, cfg)
    vp = ptr(vprev) if vprev is not None else ctypes.c_void_p(0)
    N.call("pyftle_advect", ctypes.byref(field.desc), ctypes.byref(opts), ptr(positions),
           dtype_code(positions.dtype), positions.shape[0], vp, stream_handle(stream))
    return positions


def ftle_from_groups(positions: torch.Tensor, initial_positions: torch.Tensor, map_period: float,
                     out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """Jacobian + FP64 FTLE per seed from final and initial block-ordered positions."""
    dim = positions.shape[1]
    n = positions.shape[0] // (2 * dim)
    if initial_positions.dtype != positions.dtype:
        initial_positions = initial_positions.to(positions.dtype)
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=positions.device)
    N.call("pyftle_ftle_from_groups", ptr(positions.contiguous()), ptr(initial_positions.contiguous()),
           dtype_code(positions.dtype), dim, n, float(map_period), ptr(out), stream_handle(stream))
    return out
