"""MATLAB ``.mat`` readers (file_readers.py:35-182), same keys and layouts.

* ``read_velocity``: ``velocity_x``, ``velocity_y`` (, ``velocity_z``),
  each raveled, stacked to ``(dim, N)``;
* ``read_coordinate``: ``coordinate_x``, ... likewise;
* ``read_seed_particles_coordinates``: blocks ``left, right, top, bottom``
  (, ``front, back``) concatenated along axis 0 into ``NeighboringParticles``.

``Prefetcher`` reads the files of a window ahead of their use on a few I/O
threads: file reads and numpy copies release the GIL, so reading snapshot k+1..
overlaps packing and launching snapshot k.
"""

from __future__ import annotations

import threading
from concurrent.futures import Future, ThreadPoolExecutor
from typing import Callable, Sequence

import numpy as np
from scipy.io import loadmat

from .particles import NeighboringParticles


def read_velocity(file_path: str) -> np.ndarray:
    data = loadmat(file_path)
    if "velocity_x" not in data or "velocity_y" not in data:
        raise ValueError(
            "The MATLAB file does not contain the expected keys 'velocity_x' and"
            "'velocity_y'."
        )
    vx = np.asarray(data["velocity_x"]).ravel()
    vy = np.asarray(data["velocity_y"]).ravel()
    if "velocity_z" in data:
        return np.stack((vx, vy, np.asarray(data["velocity_z"]).ravel()))
    return np.stack((vx, vy))


def read_coordinate(file_path: str) -> np.ndarray:
    data = loadmat(file_path)
    if "coordinate_x" not in data or "coordinate_y" not in data:
        raise ValueError(
            "The MATLAB file does not contain the expected keys 'coordinate_x' and"
            "'coordinate_y'."
        )
    cx = np.asarray(data["coordinate_x"]).ravel()
    cy = np.asarray(data["coordinate_y"]).ravel()
    if "coordinate_z" in data:
        return np.stack((cx, cy, np.asarray(data["coordinate_z"]).ravel()))
    return np.stack((cx, cy))


def read_seed_particles_coordinates(file_path: str) -> NeighboringParticles:
    data = loadmat(file_path)
    if "left" not in data or "right" not in data or "top" not in data or "bottom" not in data:
        raise ValueError(
            "The MATLAB file must contain at least "
            "'left', 'right', 'top' and 'bottom' keys. ('front' and 'back' for 3D)"
        )
    blocks = [np.asarray(data[k]) for k in ("left", "right", "top", "bottom")]
    if "front" in data and "back" in data:
        blocks += [np.asarray(data["front"]), np.asarray(data["back"])]
    return NeighboringParticles(positions=np.concatenate(blocks, axis=0))


_POOL = None
_POOL_LOCK = threading.Lock()
IO_THREADS = 4


def io_pool() -> ThreadPoolExecutor:
    global _POOL
    with _POOL_LOCK:
        if _POOL is None:
            _POOL = ThreadPoolExecutor(max_workers=IO_THREADS, thread_name_prefix="pyftle-io")
        return _POOL


class Prefetcher:
    """``get(i)`` returns ``reader(paths[i])``, keeping up to ``depth`` later files in flight.

    Results are handed out once; a read that raised re-raises in ``get``.
    """

    def __init__(self, paths: Sequence[str], reader: Callable = None, depth: int = 4):
        self.paths = list(paths)
        self.reader = reader or read_velocity
        self.depth = max(0, int(depth))
        self._fut: dict[int, Future] = {}

    def _submit(self, i: int) -> None:
        if 0 <= i < len(self.paths) and i not in self._fut:
            self._fut[i] = io_pool().submit(self.reader, self.paths[i])

    def get(self, i: int):
        if not 0 <= i < len(self.paths):
            raise IndexError(f"file index {i} out of range for {len(self.paths)} files")
        self._submit(i)
        for j in range(i + 1, i + 1 + self.depth):
            self._submit(j)
        return self._fut.pop(i).result()

    def __iter__(self):
        for i in range(len(self.paths)):
            yield self.get(i)
