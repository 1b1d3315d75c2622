"""``NeighboringParticles`` with pyFTLE's layout (particles.py:8-237).

Positions are ``(nb*N, dim)`` in neighbour blocks ``[left, right, top,
bottom(, front, back)]`` (file_readers.py:175-180; particles.py:170,191,212),
``nb = 2*dim`` (particles.py:76).  ``positions`` may be a NumPy array (as in
the reference) or a device tensor; the initial deltas and centroid are frozen
at construction (particles.py:72-93).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch


def _split(positions, num_neighbors):
    n = positions.shape[0] // num_neighbors
    return [positions[i * n:(i + 1) * n] for i in range(num_neighbors)]


def compute_delta_right_left(positions, num_neighbors):
    """particles.py:153-171: ``right - left`` (block 1 - block 0)."""
    left, right, *_ = _split(positions, num_neighbors)
    return right - left


def compute_delta_top_bottom(positions, num_neighbors):
    """particles.py:174-192: ``top - bottom`` (block 2 - block 3)."""
    _, _, top, bottom, *_ = _split(positions, num_neighbors)
    return top - bottom


def compute_delta_front_back(positions, num_neighbors):
    """particles.py:195-213: ``front - back`` (block 4 - block 5)."""
    *_, front, back = _split(positions, num_neighbors)
    return front - back


def compute_centroid(positions, num_neighbors):
    """particles.py:216-237: mean of the neighbour blocks."""
    parts = _split(positions, num_neighbors)
    if isinstance(positions, torch.Tensor):
        return torch.stack(parts).mean(dim=0)
    return np.mean(parts, axis=0)


@dataclass
class NeighboringParticles:
    """Groups of 4 (2D) or 6 (3D) neighbouring particles (particles.py:8-150)."""

    positions: object  # (4N, 2) or (6N, 3), numpy or torch

    initial_delta_top_bottom: object = field(init=False)
    initial_delta_right_left: object = field(init=False)
    initial_delta_front_back: object = field(init=False)
    initial_centroid: object = field(init=False)
    num_neighbors: int = field(init=False)

    def __post_init__(self) -> None:
        assert self.positions.shape[0] % 4 == 0 or self.positions.shape[0] % 6 == 0, (
            "positions.shape[0] must be multiple of 4 or 6"
        )
        self.num_neighbors = self.positions.shape[1] * 2
        self.initial_delta_top_bottom = compute_delta_top_bottom(self.positions, self.num_neighbors)
        self.initial_delta_right_left = compute_delta_right_left(self.positions, self.num_neighbors)
        if self.positions.shape[1] == 3:
            self.initial_delta_front_back = compute_delta_front_back(self.positions, self.num_neighbors)
        else:
            self.initial_delta_front_back = np.zeros(0)
        self.initial_centroid = compute_centroid(self.positions, self.num_neighbors)
        # Initial positions kept for the device Jacobian (same data the
        # initial deltas are made of).
        if isinstance(self.positions, torch.Tensor):
            self.initial_positions = self.positions.clone()
        else:
            self.initial_positions = np.array(self.positions, copy=True)

    def __len__(self) -> int:
        return self.positions.shape[0] // self.num_neighbors

    @property
    def delta_top_bottom(self):
        return compute_delta_top_bottom(self.positions, self.num_neighbors)

    @property
    def delta_right_left(self):
        return compute_delta_right_left(self.positions, self.num_neighbors)

    @property
    def delta_front_back(self):
        return compute_delta_front_back(self.positions, self.num_neighbors)

    @property
    def centroid(self):
        return compute_centroid(self.positions, self.num_neighbors)


def neighbor_positions(centers, delta: float) -> np.ndarray:
    """Block-ordered ``(2*dim*N, dim)`` neighbours at ``+-delta`` around ``centers``.

    Same construction as examples/01_Closed-form_velocity.ipynb (particle
    cell): left/right along x, top/bottom along y (front/back along z).
    """
    c = np.asarray(centers, dtype=np.float64)
    dim = c.shape[1]
    blocks = []
    for axis in range(dim):
        e = np.zeros(dim)
        e[axis] = delta
        blocks += [c - e, c + e] if axis == 0 else [c + e, c - e]
    return np.concatenate(blocks, axis=0)
