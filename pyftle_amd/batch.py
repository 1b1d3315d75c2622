"""Many FTLE windows in one persistent launch (``pyftle_batch_*``).

pyFTLE computes one FTLE field per sliding window, one window per worker
process (parallel.py:224-228, __main__.py:111-166, api.py:125-155).  Here a
batch of windows -- each with its own grid, snapshot ring, seeds, window
length, step and time origin -- is cut into work items (field, block of
seeds), ordered longest-first, and one persistent kernel pulls items from a
device counter until the list is empty.  With ``use_graph`` the counter reset
and the persistent kernel are captured once into a hipGraph and every
``run()`` replays it.

``autotune_batch`` picks seeds per item, persistent work-groups per CU and the
item order by timing graph replays.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field as dc_field

import torch

from . import _native as N
from .field import DeviceField, dtype_code, stream_handle
from .window import METHODS, TIME_MODES


class BatchField(ctypes.Structure):
    _fields_ = [
        ("field", N.Field),
        ("positions_in", ctypes.c_void_p),
        ("positions_out", ctypes.c_void_p),
        ("ftle_out", ctypes.c_void_p),
        ("n_groups", ctypes.c_int64),
        ("n_steps", ctypes.c_int32),
        ("first_snapshot", ctypes.c_int32),
        ("t_start", ctypes.c_double),
        ("h", ctypes.c_double),
        ("map_period", ctypes.c_double),
        ("seed_order", ctypes.c_void_p),
    ]


class BatchOpts(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("time_mode", ctypes.c_int32),
        ("pos_dtype", ctypes.c_int32),
        ("groups_per_block", ctypes.c_int32),
        ("lane_order", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("longest_first", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
    ]


_SIGS = {
    "pyftle_batch_create": [ctypes.POINTER(BatchField), ctypes.c_int32, ctypes.POINTER(BatchOpts),
                            ctypes.POINTER(ctypes.c_void_p)],
    "pyftle_batch_launch": [ctypes.c_void_p, ctypes.c_void_p],
    "pyftle_batch_info": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32),
                          ctypes.POINTER(ctypes.c_int32)],
    "pyftle_batch_destroy": [ctypes.c_void_p],
    "pyftle_batch_count_loads": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "pyftle_batch_run_noalloc": [ctypes.POINTER(BatchField), ctypes.c_int32, ctypes.POINTER(BatchOpts),
                                 ctypes.c_void_p],
}

NOALLOC_MAX_FIELDS = 256  # PYFTLE_BATCH_NOALLOC_MAX_FIELDS


@dataclass
class BatchConfig:
    groups_per_block: int = 64
    lane_order: int = N.ORDER_SEED_MAJOR
    blocks_per_cu: int = 0
    longest_first: int = 1
    use_graph: int = 1


@dataclass
class Window:
    """One FTLE window of a batch."""

    field: DeviceField
    positions: torch.Tensor          # block-ordered (2*dim*N, dim) device tensor
    n_steps: int
    h: float
    map_period: float
    t_start: float = 0.0
    first_snapshot: int = 0
    positions_out: torch.Tensor | None = None
    ftle: torch.Tensor | None = dc_field(default=None)
    order: torch.Tensor | None = None  # int32 device permutation of the seed groups (window.seed_order)


def _describe(windows: list[Window], method, time_mode, cfg: BatchConfig):
    """Checks the windows and fills the ``pyftle_batch_field`` array and options;
    allocates each window's missing FTLE output."""
    if not windows:
        raise ValueError("a batch needs at least one window")
    pdt = windows[0].positions.dtype
    arr = (BatchField * len(windows))()
    for i, w in enumerate(windows):
        dim = w.field.geom.dim
        p = w.positions
        if p.dtype != pdt or not p.is_cuda or not p.is_contiguous():
            raise ValueError("batch positions must be contiguous device tensors of one dtype")
        if p.dim() != 2 or p.shape[1] != dim or p.shape[0] % (2 * dim) != 0:
            raise ValueError(f"window {i}: positions must have shape (2*{dim}*N, {dim}), got {tuple(p.shape)}")
        n = p.shape[0] // (2 * dim)
        po = w.positions_out
        if po is not None and (po.shape != p.shape or po.dtype != p.dtype or not po.is_cuda
                               or not po.is_contiguous()):
            raise ValueError(f"window {i}: positions_out must be a contiguous device tensor like positions")
        if w.ftle is None:
            w.ftle = torch.empty(n, dtype=torch.float64, device=p.device)
        elif (w.ftle.dtype != torch.float64 or w.ftle.numel() != n or not w.ftle.is_cuda
              or not w.ftle.is_contiguous()):
            raise ValueError(f"window {i}: ftle must be a contiguous float64 device tensor of {n} values")
        bf = arr[i]
        bf.field = w.field.desc
        bf.positions_in = w.positions.data_ptr()
        bf.positions_out = w.positions_out.data_ptr() if w.positions_out is not None else None
        bf.ftle_out = w.ftle.data_ptr()
        bf.n_groups = n
        bf.n_steps = int(w.n_steps)
        bf.first_snapshot = int(w.first_snapshot)
        bf.t_start = float(w.t_start)
        bf.h = float(w.h)
        bf.map_period = float(w.map_period)
        if w.order is not None:
            if w.order.dtype != torch.int32 or not w.order.is_cuda or w.order.numel() != n:
                raise ValueError(f"window {i}: order must be an int32 device permutation of {n} groups")
            w.order = w.order.contiguous()
            bf.seed_order = w.order.data_ptr()
    o = BatchOpts()
    o.method = METHODS[method] if isinstance(method, str) else int(method)
    o.time_mode = TIME_MODES[time_mode] if isinstance(time_mode, str) else int(time_mode)
    o.pos_dtype = dtype_code(pdt)
    o.groups_per_block = cfg.groups_per_block
    o.lane_order = cfg.lane_order
    o.blocks_per_cu = cfg.blocks_per_cu
    o.longest_first = cfg.longest_first
    o.use_graph = cfg.use_graph
    return arr, o


class BatchPlan:
    """A captured batch: ``run()`` recomputes every window's FTLE in one launch.

    The device pointers of every window (positions, outputs, snapshot rings) are
    bound at construction and replayed by each ``run()``; the plan keeps the
    windows alive.  Runs of one plan share its work counter, so two runs of the
    same plan must not overlap on different streams.  ``close()`` waits for the
    device before freeing the plan."""

    def __init__(self, windows: list[Window], method="rk4", time_mode="linear",
                 cfg: BatchConfig | None = None):
        N.extra_signatures(_SIGS)
        self.windows = windows
        self.cfg = cfg or BatchConfig()
        if not windows:
            raise ValueError("a batch needs at least one window")
        arr, o = _describe(windows, method, time_mode, self.cfg)
        h = ctypes.c_void_p()
        N.call("pyftle_batch_create", arr, len(windows), ctypes.byref(o), ctypes.byref(h))
        self._h = h
        self._arr = arr

    def info(self):
        n = ctypes.c_int64()
        g = ctypes.c_int32()
        t = ctypes.c_int32()
        N.call("pyftle_batch_info", self._h, ctypes.byref(n), ctypes.byref(g), ctypes.byref(t))
        return {"items": n.value, "grid": g.value, "threads": t.value}

    def run(self, stream=None) -> list[torch.Tensor]:
        N.call("pyftle_batch_launch", self._h, stream_handle(stream))
        return [w.ftle for w in self.windows]

    def count_loads(self, stream=None) -> int:
        """Runs the batch once with a counting copy of the persistent kernel (RK4 on
        per-cell records; the same FTLE fields) and returns the wave-level 16-byte record
        loads it issued (``pyftle_batch_count_loads``).  Waits for the device."""
        loads = torch.zeros(1, dtype=torch.int64, device=self.windows[0].positions.device)
        N.call("pyftle_batch_count_loads", self._h, loads.data_ptr(), stream_handle(stream))
        return int(loads.item())

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            N.load().pyftle_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_batch_noalloc(windows: list[Window], method="rk4", time_mode="linear",
                      cfg: BatchConfig | None = None, stream=None) -> list[torch.Tensor]:
    """One batch through ``pyftle_batch_run_noalloc``: the same FTLE fields as
    ``BatchPlan(windows, ...).run()`` with no device allocation by the library (its
    descriptors and work counter live in a static per-GPU array).  Each window's
    ``ftle`` must already be allocated, or is allocated here by torch before the
    call.  Calls that overlap on one GPU race on that array."""
    N.extra_signatures(_SIGS)
    if len(windows) > NOALLOC_MAX_FIELDS:
        raise ValueError(f"at most {NOALLOC_MAX_FIELDS} windows per allocation-free batch")
    arr, o = _describe(windows, method, time_mode, cfg or BatchConfig())
    N.call("pyftle_batch_run_noalloc", arr, len(windows), ctypes.byref(o), stream_handle(stream))
    return [w.ftle for w in windows]


def autotune_batch(windows, method="rk4", time_mode="linear", repeats=2):
    """Time graph replays for each candidate configuration; returns (best, table)."""
    dim = windows[0].field.geom.dim
    table = []
    for g in (32, 64, 128):
        if g * 2 * dim > 512:
            continue
        for bpc in (0, 2, 4):
            for lf in (1, 0):
                cfg = BatchConfig(groups_per_block=g, blocks_per_cu=bpc, longest_first=lf)
                plan = BatchPlan(windows, method, time_mode, cfg)
                plan.run()
                torch.cuda.synchronize()
                best = float("inf")
                for _ in range(repeats):
                    a = torch.cuda.Event(enable_timing=True)
                    b = torch.cuda.Event(enable_timing=True)
                    a.record()
                    plan.run()
                    b.record()
                    torch.cuda.synchronize()
                    best = min(best, a.elapsed_time(b))
                plan.close()
                table.append((best, cfg))
    table.sort(key=lambda r: r[0])
    return table[0][1], [(ms, vars(c)) for ms, c in table]
