"""Notebook API ``AnalyticalSolver`` (api.py:15-224) on the device.

Same constructor and result: ``num_snapshots = int(T/|dt|) + 1`` (api.py:105),
``num_ftles`` windows starting at ``k*dt`` (api.py:125-155), one integrator
shared by every window (api.py:112,177-183), results stacked to
``(num_ftles, N)`` (api.py:224) or a single ``(N,)`` field.

Windows run one after another on the GPU instead of in a process pool; each
starts from the caller's original particle positions, as every pickled task
of the reference's pool does.  The analytical velocity is evaluated as in the
reference, at its frozen time 0 (interpolate.py:788,812) at every integrator
stage, on device tensors: ``velocity_fn(t, points)`` receives a CUDA tensor
and returns one (torch operations).  ``grid=(lower, upper, shape)`` opts in to
sampling the function once onto a device grid of that box and resolution
(``precision`` sets its storage), read by the fused window kernel; a particle
leaving that box raises ``ValueError``.
"""

from __future__ import annotations

from datetime import datetime
from typing import Callable, Optional

import numpy as np
import torch

from .data_source import AnalyticalBatchSource
from .ftle_solver import FTLESolver
from .integrate import create_integrator
from .interpolate import create_interpolator
from .particles import NeighboringParticles


class AnalyticalSolver:
    def __init__(self, velocity_fn: Callable, particles: NeighboringParticles, timestep: float,
                 flow_map_period: float, num_ftles: int, integrator_name: str, num_processes: int = 1,
                 save_output: bool = False, output_format: str = "vtk",
                 output_dir_name: Optional[str] = None, grid=None, precision: str = "float32"):
        self.velocity_fn = velocity_fn
        self.particles = particles
        self.timestep = timestep
        self.flow_map_period = flow_map_period
        self.num_ftles = num_ftles
        self.num_snapshots = int(flow_map_period / abs(timestep)) + 1
        self.writer = None
        self.num_processes = num_processes
        interpolator = create_interpolator("analytical", velocity_fn=velocity_fn)
        if grid is not None:
            lower, upper, shape = grid
            interpolator.set_grid(lower, upper, shape, precision=precision)
        self.integrator = create_integrator(integrator_name, interpolator)
        if self.timestep < 0:
            print("Running backward-time FTLE")
        else:
            print("Running forward-time FTLE")
        if save_output:
            from .file_writers import create_writer

            if output_dir_name is None:
                output_dir_name = datetime.now().strftime("run-%Y-%m-%d-%Hh-%Mm-%Ss")
            self.writer = create_writer(output_format, output_dir_name)

    def _create_batches(self):
        start_times = np.arange(self.num_ftles) * self.timestep
        offsets = np.arange(self.num_snapshots) * self.timestep
        time_batches = start_times[:, None] + offsets
        return [
            AnalyticalBatchSource(self.velocity_fn, self._fresh_particles(), self.timestep, time_batches[i])
            for i in range(self.num_ftles)
        ]

    def _fresh_particles(self) -> NeighboringParticles:
        p = self.particles.positions
        copy = p.clone() if isinstance(p, torch.Tensor) else np.array(p, copy=True)
        return NeighboringParticles(positions=copy)

    def run(self):
        results = []
        for batch in self._create_batches():
            try:
                solver = FTLESolver(batch, integrator=self.integrator, output_writer=self.writer)
                results.append(solver.run())
            except Exception as e:  # parallel.py:236-269: stop on the first failure
                raise RuntimeError("One or more FTLE batches failed.") from e
        if self.writer is not None:
            return None
        if not results:
            raise RuntimeError("No FTLE fields were returned (all results were None).")
        if len(results) == 1:
            return results[0]
        if isinstance(results[0], torch.Tensor):
            return torch.stack(results, dim=0)
        return np.stack(results, axis=0)
