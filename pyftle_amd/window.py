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


def column_box(x: torch.Tensor):
    """Per-column min and max of x (n, d), NaN where a column holds one.  A reduction over a
    tall, narrow tensor is slow along dim 0 (16.8M x 3: 9.7 ms), so the rows are reduced 64 at a
    time as one row of 64 d values (0.17 ms), the 64 partial results folded afterwards."""
    n, d = x.shape
    if n == 0:
        return torch.aminmax(x, dim=0)  # torch's own error for an empty reduction
    b = 64
    nb = n // b
    parts = []
    if nb:
        lo, hi = torch.aminmax(x[:nb * b].reshape(nb, b * d), dim=0)
        parts.append((lo.view(b, d).amin(0), hi.view(b, d).amax(0)))
    if n > nb * b:
        parts.append(torch.aminmax(x[nb * b:], dim=0))
    return torch.stack([q[0] for q in parts]).amin(0), torch.stack([q[1] for q in parts]).amax(0)


def seed_order(field: DeviceField, positions: torch.Tensor, nb: int) -> torch.Tensor:
    """Processing order of the seed groups of a block-ordered ``(nb*N, dim)`` array:
    a stable sort by the Morton (Z-order) code of the grid cell holding each group's
    centroid (``pyftle_seed_morton_codes``).  Spatially close seeds then share waves and
    their gathers share cache lines; results do not depend on the order."""
    dim = field.geom.dim
    n = positions.shape[0] // nb
    codes = torch.empty(n, dtype=torch.int64, device=positions.device)
    if field.interp in (N.INTERP_SCATTER_LINEAR, N.INTERP_SCATTER_NEAREST, N.INTERP_SCATTER_CUBIC):
        lo_t, hi_t = column_box(positions)
        lo = [float(v) for v in lo_t.tolist()]
        ext = [max(float(h - l), 1e-300) for l, h in zip(lo, hi_t.tolist())]
        inv = [1024.0 / e for e in ext]
    else:
        lo = list(field.geom.origin)
        inv = [1.0 / sp for sp in field.geom.spacing]
    lo3 = (ctypes.c_double * 3)(*(lo + [0.0] * (3 - dim)))
    inv3 = (ctypes.c_double * 3)(*(inv + [0.0] * (3 - dim)))
    N.call("pyftle_seed_morton_codes", ptr(positions), dtype_code(positions.dtype), n, nb, dim, lo3, inv3,
           ptr(codes), stream_handle())
    return torch.argsort(codes, stable=True).to(torch.int32)


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


def tuned_config(field: DeviceField, n_groups: int, time_mode) -> LaunchConfig:
    """Launch shape from the tuning cache for this field's shape class (dimension, time
    treatment, precision, seed count, record layout), the defaults when it was never tuned.
    The shape changes only the speed: every seed group is independent."""
    from .tuning import best_config, shape_class

    tm = time_mode if isinstance(time_mode, str) else {v: k for k, v in TIME_MODES.items()}.get(time_mode, "")
    prec = "f32" if field.dtype == torch.float32 else "f64"
    cell = field.layout in (N.LAYOUT_CELL, N.LAYOUT_CELL_PAIR)
    return best_config(shape_class(field.geom.dim, n_groups, tm, prec, records="cell" if cell else "node"))


def advect_ftle(field: DeviceField, positions: torch.Tensor, *, method="rk4", time_mode="frozen",
                n_steps: int, h: float, map_period: float, first: int = 0, t_start: float = 0.0,
                positions_out: torch.Tensor | None = None, ftle_out: torch.Tensor | None = None,
                cfg: LaunchConfig | None = None, stream=None, order=None,
                count_loads: torch.Tensor | None = None) -> torch.Tensor:
    """Advect block-ordered seed groups and return the FP64 FTLE per seed.

    ``order``: ``None``, ``"morton"`` (``seed_order``) or an int32 device permutation of the
    seed groups giving the order in which they are processed (results are the same).
    ``count_loads``: a one-element int64 device tensor; the launch then uses the counting
    copy of the kernel (``pyftle_advect_ftle_count_loads``, same results) and adds the
    wave-level record loads it issues (RK4 on per-cell FP32 records only)."""
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
    if cfg is None:
        cfg = tuned_config(field, n_groups, time_mode)
    opts = make_opts(method, time_mode, n_steps, h, first, t_start, False, cfg)
    if isinstance(order, str):
        if order != "morton":
            raise ValueError(f"unknown seed order {order!r}")
        order = seed_order(field, positions, nb)
    if order is not None:
        if order.dtype != torch.int32 or not order.is_cuda or order.numel() != n_groups:
            raise ValueError("order must be an int32 device permutation of the seed groups")
        order = order.contiguous()
        opts.seed_order = order.data_ptr()
    pout = ptr(positions_out) if positions_out is not None else ctypes.c_void_p(0)
    if count_loads is not None:
        if count_loads.dtype != torch.int64 or not count_loads.is_cuda or count_loads.numel() != 1:
            raise ValueError("count_loads must be a one-element int64 device tensor")
        N.call("pyftle_advect_ftle_count_loads", ctypes.byref(field.desc), ctypes.byref(opts), ptr(positions),
               dtype_code(positions.dtype), n_groups, pout, float(map_period), ptr(ftle_out),
               ptr(count_loads), stream_handle(stream))
        return ftle_out
    N.call("pyftle_advect_ftle", ctypes.byref(field.desc), ctypes.byref(opts), ptr(positions),
           dtype_code(positions.dtype), n_groups, pout, float(map_period), ptr(ftle_out),
           stream_handle(stream))
    return ftle_out


def probe_load_rate(blocks: int = 2048, iters: int = 8192, repeats: int = 5, stream=None) -> dict:
    """Peak rate of 16-byte-per-lane wave loads through the vector-memory path
    (``pyftle_probe_load_rate``: coalesced, L1-resident), best of ``repeats`` launches
    timed with HIP events on the launch stream."""
    buf = torch.zeros(4096, dtype=torch.float32, device="cuda")
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
    st = stream or torch.cuda.current_stream()
    N.call("pyftle_probe_load_rate", ptr(buf), int(blocks), int(iters), ptr(out), stream_handle(st))
    best = float("inf")
    for _ in range(repeats):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(st)
        N.call("pyftle_probe_load_rate", ptr(buf), int(blocks), int(iters), ptr(out), stream_handle(st))
        b.record(st)
        b.synchronize()
        best = min(best, a.elapsed_time(b) * 1e-3)
    wave_loads = blocks * 4 * iters
    return {"wave_loads": wave_loads, "seconds": best, "wave_loads_per_s": wave_loads / best,
            "bytes_per_s": wave_loads * 1024 / best}


ISSUE_MODES = {0: "mfma", 1: "mul", 2: "exp2", 3: "rcp", 4: "mfma+2mul one wave",
               5: "mfma wave | mul wave", 6: "mfma wave | exp2 wave", 7: "mfma wave | rcp wave"}


def probe_issue_rates(iters: int = 20000, stream=None) -> dict:
    """Cycles per instruction of the fused MLP kernels' mix (``pyftle_probe_issue_rates``),
    one work-group per CU: per mode, the mean s_memtime cycles per slot-instruction of the
    MFMA wave (``first``) and, for the paired modes, of the VALU wave sharing its SIMD
    (``second``).  A slot holds one MFMA or two VALU instructions."""
    blocks = torch.cuda.get_device_properties(0).multi_processor_count
    cycles = torch.zeros(2 * blocks, dtype=torch.int64, device="cuda")
    out = torch.empty(blocks * 512, dtype=torch.float32, device="cuda")
    st = stream or torch.cuda.current_stream()
    res = {}
    N.call("pyftle_probe_issue_rates", 0, blocks, 20 * iters, ptr(cycles), ptr(out), stream_handle(st))  # clock up
    for mode, name in ISSUE_MODES.items():
        N.call("pyftle_probe_issue_rates", mode, blocks, iters, ptr(cycles), ptr(out), stream_handle(st))
        st.synchronize()
        c = cycles.view(blocks, 2).double().mean(0).cpu().numpy() / iters
        per_mfma = c[0] / 8
        per_valu = (c[0] if mode in (1, 2, 3) else c[1]) / 16
        row = {"cycles_per_iteration": [float(c[0]), float(c[1])]}
        if mode in (0, 4, 5, 6, 7):
            row["cycles_per_mfma"] = float(per_mfma)
        if mode in (1, 2, 3, 5, 6, 7):
            row["cycles_per_valu"] = float(per_valu)
        res[name] = row
    return res


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
