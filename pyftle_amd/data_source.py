"""Window feeds (data_source.py:29-285): the ``BatchSource`` protocol.

* ``FileBatchSource``: one ``.mat`` velocity file per step (read ahead by a
  ``Prefetcher``), coordinates read on step 0 and re-read only when the
  coordinate files differ (data_source.py:86-211);
* ``AnalyticalBatchSource``: closed-form velocity, no per-step update
  (data_source.py:214-285);
* ``ArrayBatchSource`` (new): snapshots already in memory (host arrays or
  device tensors) in the reference's ``(dim, Ngrid)`` layout.
"""

from __future__ import annotations

from pathlib import Path
from typing import Callable, List, Protocol, Sequence, Tuple

from .file_readers import Prefetcher, read_coordinate, read_seed_particles_coordinates, read_velocity
from .particles import NeighboringParticles


class BatchSource(Protocol):
    @property
    def timestep(self) -> float: ...

    @property
    def num_steps(self) -> int: ...

    @property
    def id(self) -> str: ...

    def get_particles(self) -> NeighboringParticles: ...

    def update_interpolator(self, interpolator, step_index: int) -> None: ...


class FileBatchSource:
    """data_source.py:86-211."""

    def __init__(self, snapshot_files: List[str], coordinate_files: List[str], particle_file: str,
                 snapshot_timestep: float, flow_map_period: int | float):
        self.snapshot_files = snapshot_files
        self.coordinate_files = coordinate_files
        self.particle_file = particle_file
        self.snapshot_timestep = snapshot_timestep
        self.flow_map_period = flow_map_period
        self._n = len(snapshot_files)
        self._id = f"{Path(self.snapshot_files[0]).stem}"
        self.reuse_coordinates = len(set(coordinate_files)) == 1
        self._velocities = Prefetcher(snapshot_files, read_velocity)  # reads ahead on I/O threads

    @property
    def id(self) -> str:
        return self._id

    @property
    def num_steps(self) -> int:
        return self._n

    @property
    def timestep(self) -> float:
        return self.snapshot_timestep

    def get_particles(self) -> NeighboringParticles:
        return read_seed_particles_coordinates(self.particle_file)

    def get_data_for_step(self, step_index: int) -> Tuple:
        velocities = self._velocities.get(step_index)
        coordinates = None
        if step_index == 0 or not self.reuse_coordinates:
            coordinates = read_coordinate(self.coordinate_files[step_index])
        return velocities, coordinates

    def update_interpolator(self, interpolator, step_index: int) -> None:
        velocities, coordinates = self.get_data_for_step(step_index)
        interpolator.update(velocities, coordinates)


class AnalyticalBatchSource:
    """data_source.py:214-285."""

    def __init__(self, velocity_fn: Callable, particles: NeighboringParticles, timestep: float, times):
        self.velocity_fn = velocity_fn
        self.particles = particles
        self._timestep = timestep
        self.times = times

    @property
    def id(self) -> str:
        return f"{self.times[0]:06f}"

    @property
    def num_steps(self) -> int:
        return len(self.times)

    @property
    def timestep(self) -> float:
        return self._timestep

    def get_particles(self) -> NeighboringParticles:
        return self.particles

    def update_interpolator(self, interpolator, step_index: int) -> None:
        pass


class ArrayBatchSource:
    """In-memory window: ``velocities[i]`` is the ``(dim, Ngrid)`` snapshot of step i."""

    def __init__(self, velocities: Sequence, coordinates, particles: NeighboringParticles,
                 timestep: float, id: str = "window"):
        self.velocities = velocities
        self.coordinates = coordinates
        self.particles = particles
        self._timestep = timestep
        self._id = id

    @property
    def id(self) -> str:
        return self._id

    @property
    def num_steps(self) -> int:
        return len(self.velocities)

    @property
    def timestep(self) -> float:
        return self._timestep

    def get_particles(self) -> NeighboringParticles:
        return self.particles

    def update_interpolator(self, interpolator, step_index: int) -> None:
        interpolator.update(self.velocities[step_index], self.coordinates if step_index == 0 else None)
