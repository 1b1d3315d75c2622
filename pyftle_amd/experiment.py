"""File-driven experiments: the reference's configuration keys and window construction.

* ``ExperimentConfig`` / ``load_config``: the keys of ``config.yaml`` and
  ``hyperparameters.py:96-255`` (``experiment_name``, ``list_*_files``,
  ``snapshot_timestep``, ``flow_map_period``, ``integrator``, ``interpolator``,
  ``num_processes``, ``output_format``, ``flow_grid_shape``,
  ``particles_grid_shape``), with ``parse_tuple``'s rules for grid shapes.
* ``FileFTLEManager``: ``MultipleFTLEProcessManager`` (``__main__.py:55-166``).
  File lists are read one path per line (``file_utils.py:96-98``); a negative
  timestep reverses the three lists (``__main__.py:90-108``); windows are
  ``p = int(T / |dt|) + 1`` consecutive snapshots starting at every index
  ``i < n - p + 1``, with coordinate files cycled over ``n + p - 1`` entries
  and particle files over ``n`` (``__main__.py:111-166``).  Results come back in
  window order, or go to the writer as ``ftle_<id>`` (``ftle_solver.py:112-119``).

Instead of one process per window (``parallel.py:171-271``) the manager runs on
one GPU.  ``batched=True`` (the default when every window shares one grid and
the ``grid`` interpolator) reads each snapshot file once into a single device
ring and computes every window in one persistent launch (``BatchPlan``), each
window reading its own snapshots frozen per step, as ``FTLESolver`` does; the
FTLE fields are bit-identical to running ``FTLESolver`` window by window.
"""

from __future__ import annotations

import itertools
import os
from dataclasses import dataclass, field
from typing import List, Optional

import torch

from . import _native as N
from .batch import BatchConfig, BatchPlan, Window
from .data_source import FileBatchSource
from .field import DeviceField, GridGeometry, to_device, to_host
from .file_readers import Prefetcher, read_coordinate, read_seed_particles_coordinates, read_velocity
from .ftle_solver import FTLESolver
from .integrate import create_integrator
from .interpolate import create_interpolator

_REQUIRED = ("experiment_name", "list_velocity_files", "list_coordinate_files", "list_particle_files",
             "snapshot_timestep", "flow_map_period")


def parse_tuple(value) -> tuple:
    """``"10,20,30"`` -> ``(10, 20, 30)``; 2 or 3 positive integers, a trailing 1 dropped."""
    if isinstance(value, (tuple, list)):
        vals = [int(v) for v in value]
    else:
        vals = [int(v) for v in str(value).split(",")]
    if len(vals) not in (2, 3):
        raise ValueError("Tuple must contain either 2 or 3 elements.")
    if not all(v > 0 for v in vals):
        raise ValueError("All elements must be positive integers.")
    if vals[-1] == 1:
        vals.pop()
    return tuple(vals)


@dataclass
class ExperimentConfig:
    experiment_name: str
    list_velocity_files: str
    list_coordinate_files: str
    list_particle_files: str
    snapshot_timestep: float
    flow_map_period: float
    integrator: Optional[str] = None
    interpolator: Optional[str] = None
    num_processes: int = 1
    output_format: Optional[str] = None
    flow_grid_shape: Optional[tuple] = None
    particles_grid_shape: Optional[tuple] = None
    extra: dict = field(default_factory=dict)  # keys this build does not know


def load_config(path: str | None = None, **overrides) -> ExperimentConfig:
    """Read the YAML keys of ``config.yaml`` (safe loader); keyword overrides win,
    as command-line flags do over the config file."""
    import yaml

    raw = {}
    if path is not None:
        with open(path) as f:
            raw = yaml.safe_load(f) or {}
    raw.update({k: v for k, v in overrides.items() if v is not None})
    missing = [k for k in _REQUIRED if k not in raw]
    if missing:
        raise ValueError(f"missing required configuration keys: {', '.join(missing)}")
    known = set(ExperimentConfig.__dataclass_fields__) - {"extra"}
    kw = {k: v for k, v in raw.items() if k in known}
    for k in ("flow_grid_shape", "particles_grid_shape"):
        if kw.get(k) not in (None, ""):
            kw[k] = parse_tuple(kw[k])
        else:
            kw[k] = None
    kw["snapshot_timestep"] = float(kw["snapshot_timestep"])
    kw["flow_map_period"] = float(kw["flow_map_period"])
    kw["num_processes"] = int(kw.get("num_processes", 1))
    return ExperimentConfig(**kw, extra={k: v for k, v in raw.items() if k not in known})


def get_files_list(file_path: str) -> List[str]:
    """One path per line (the first column of a header-less CSV)."""
    if not os.path.exists(file_path):
        raise FileNotFoundError(f"The file {file_path} does not exist.")
    with open(file_path) as f:
        return [line.split(",")[0].strip() for line in f if line.strip()]


class FileFTLEManager:
    def __init__(self, cfg: ExperimentConfig, precision: str = "float32", output_dir: str | None = None,
                 write: bool | None = None):
        self.cfg = cfg
        self.snapshot_files = get_files_list(cfg.list_velocity_files)
        self.coordinate_files = get_files_list(cfg.list_coordinate_files)
        self.particle_files = get_files_list(cfg.list_particle_files)
        self.timestep = cfg.snapshot_timestep
        self.flow_map_period = cfg.flow_map_period
        self.precision = precision
        interp_name = cfg.interpolator or "cubic"
        kw = {"precision": precision} if cfg.flow_grid_shape is not None else {}
        self.interpolator = create_interpolator(interp_name, cfg.flow_grid_shape, **kw)
        self.integrator = create_integrator(cfg.integrator or "euler", self.interpolator)
        self.writer = None
        if write or (write is None and cfg.output_format):
            from .file_writers import create_writer

            out = output_dir or os.path.join("outputs", cfg.experiment_name)
            self.writer = create_writer(cfg.output_format or "mat", out, cfg.particles_grid_shape)
        if self.timestep < 0:
            self.snapshot_files.reverse()
            self.coordinate_files.reverse()
            self.particle_files.reverse()

    def window_length(self) -> int:
        return int(self.flow_map_period / abs(self.timestep)) + 1

    def create_batches(self) -> list[FileBatchSource]:
        n, p = len(self.snapshot_files), self.window_length()
        coords = list(itertools.islice(itertools.cycle(self.coordinate_files), n + p - 1))
        parts = list(itertools.islice(itertools.cycle(self.particle_files), n))
        return [FileBatchSource(snapshot_files=self.snapshot_files[i:i + p], coordinate_files=coords[i:i + p],
                                particle_file=parts[i], snapshot_timestep=self.timestep, flow_map_period=p)
                for i in range(n - p + 1)]

    # -- execution -----------------------------------------------------------
    def run(self, batched: bool | None = None):
        batches = self.create_batches()
        if batched is None:
            batched = self._can_batch(batches)
        results = self._run_batched(batches) if batched else self._run_sequential(batches)
        if self.writer is not None:
            centroids = {}
            for b, r in zip(batches, results):
                key = os.path.abspath(b.particle_file)
                if key not in centroids:
                    centroids[key] = read_seed_particles_coordinates(b.particle_file).initial_centroid
                self.writer.write(f"ftle_{b.id}", r, centroids[key])
            return None
        return results

    def _run_sequential(self, batches):
        out = []
        for b in batches:
            out.append(FTLESolver(b, self.integrator).run())
        return out

    def _can_batch(self, batches) -> bool:
        if not batches or (self.cfg.interpolator or "").lower() != "grid" or self.precision != "float32":
            return False
        return len({os.path.abspath(c) for c in self.coordinate_files}) == 1

    def _run_batched(self, batches):
        shape = self.cfg.flow_grid_shape
        if shape is None:
            raise ValueError("the batched path needs flow_grid_shape (the 'grid' interpolator)")
        geom = GridGeometry.from_reference(shape, read_coordinate(self.coordinate_files[0]))
        n, p = len(self.snapshot_files), self.window_length()
        field = DeviceField(geom, dtype="float32", layout=N.LAYOUT_SINGLE, ring_slots=n, n_snapshots=n)
        # each file read once (ahead, on I/O threads) and staged through pinned buffers
        for k, vel in enumerate(Prefetcher(self.snapshot_files, read_velocity, depth=8)):
            field.pack(k, vel)
        method = {0: "euler", 1: "ab2", 2: "rk4"}[self.integrator.method]
        windows = []
        seeds = {}  # particle file -> device positions (the kernel only reads them)
        for i, b in enumerate(batches):
            key = os.path.abspath(b.particle_file)
            if key not in seeds:
                seeds[key] = to_device(read_seed_particles_coordinates(b.particle_file).positions, torch.float64)
            windows.append(Window(field=field, positions=seeds[key], n_steps=p, h=self.timestep,
                                  map_period=(p - 1) * abs(self.timestep), first_snapshot=i))
        plan = BatchPlan(windows, method=method, time_mode="frozen", cfg=BatchConfig())
        res = plan.run()
        sizes = [r.numel() for r in res]
        flat = to_host(torch.cat(res))  # one transfer for every window's field
        plan.close()
        out, o = [], 0
        for n_ in sizes:
            out.append(flat[o:o + n_])
            o += n_
        return out


__all__ = ["ExperimentConfig", "load_config", "parse_tuple", "get_files_list", "FileFTLEManager"]
