"""Launch-shape autotuner for the fused advection kernel.

The tunables are the seeds per work-group (``groups_per_block``), the lane
order (seed-major: a seed's neighbours in adjacent lanes; neighbour-major: a
wave holds one neighbour direction of consecutive seeds) and the XCD-aware
block remap.  ``autotune`` times every candidate with HIP events on a short
window (a few steps of the real field and seeds) and stores the winner per
shape class in ``tuning_cache.json`` next to this file; ``best_config`` reads
it back (defaults when a class was never tuned).
"""

from __future__ import annotations

import json
import os
import threading

from . import _native as N
from .window import LaunchConfig

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")
_lock = threading.Lock()


def shape_class(dim: int, n_seeds: int, time_mode: str, precision: str = "f32", records: str = "node") -> str:
    """Bucket by dimension, time treatment, precision, log2 of the seed count and, for the
    per-cell record layouts, the layout (their best launch shape differs)."""
    b = max(0, int(n_seeds).bit_length() - 1)
    key = f"d{dim}-{time_mode}-{precision}-s{b}"
    return key + "-cell" if records == "cell" else key


_memo = {"mtime": None, "data": {}}


def _read():
    """The cache file's entries (re-read only when the file changed)."""
    try:
        mtime = os.path.getmtime(CACHE)
    except OSError:
        return {}
    if _memo["mtime"] != mtime:
        try:
            with open(CACHE) as f:
                data = json.load(f)
        except (OSError, ValueError):
            data = {}
        _memo["mtime"], _memo["data"] = mtime, data
    return dict(_memo["data"])


def best_config(key: str) -> LaunchConfig:
    entry = _read().get(key)
    if not entry:
        return LaunchConfig()
    return LaunchConfig(groups_per_block=int(entry["groups_per_block"]),
                        lane_order=int(entry["lane_order"]), xcd_remap=int(entry["xcd_remap"]))


def candidates(dim: int):
    nb = 2 * dim
    out = []
    for g in (16, 32, 64, 128 if nb * 128 <= 512 else 85):
        if g * nb > 512:
            continue
        for order in (N.ORDER_SEED_MAJOR, N.ORDER_NEIGHBOR_MAJOR):
            for remap in (0, 1):
                out.append(LaunchConfig(groups_per_block=g, lane_order=order, xcd_remap=remap))
    return out


def autotune(run, dim: int, key: str, repeats: int = 3, save: bool = True):
    """Time ``run(cfg)`` (one launch, enqueued on the current stream) for every
    candidate and keep the fastest.  Returns ``(best, table)``."""
    import torch

    table = []
    for cfg in candidates(dim):
        run(cfg)  # warm
        torch.cuda.synchronize()
        evs = []
        for _ in range(repeats):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            run(cfg)
            b.record()
            evs.append((a, b))
        torch.cuda.synchronize()
        ms = min(a.elapsed_time(b) for a, b in evs)
        table.append((ms, cfg))
    table.sort(key=lambda r: r[0])
    best = table[0][1]
    if save:
        with _lock:
            data = _read()
            data[key] = {"groups_per_block": best.groups_per_block, "lane_order": best.lane_order,
                         "xcd_remap": best.xcd_remap, "ms": table[0][0]}
            with open(CACHE, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            _memo["mtime"] = None
    return best, [(ms, vars(c)) for ms, c in table]


def best_batch_config(key: str):
    """Cached batch configuration (``autotune_batch``), or None."""
    from .batch import BatchConfig

    entry = _read().get(key)
    if not entry:
        return None
    return BatchConfig(groups_per_block=int(entry["groups_per_block"]),
                       lane_order=int(entry.get("lane_order", 0)),
                       blocks_per_cu=int(entry.get("blocks_per_cu", 0)),
                       longest_first=int(entry.get("longest_first", 1)),
                       use_graph=int(entry.get("use_graph", 1)))


def save_batch_config(key: str, cfg, ms: float) -> None:
    with _lock:
        data = _read()
        data[key] = dict(vars(cfg), ms=ms)
        with open(CACHE, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        _memo["mtime"] = None
