"""Window driver (ftle_solver.py:15-161) on the fused HIP kernel.

``FTLESolver(source, integrator, progress_queue=None, output_writer=None)``
keeps the reference signature and result.  ``run()``:

1. ``particles = source.get_particles()``;
2. for ``i in range(num_steps)``: ``source.update_interpolator(interp, i)``
   packs snapshot ``i`` into the device ring (ftle_solver.py:99-102);
3. as soon as the ring is full (or the window is complete) the advection
   kernel takes all pending steps in one launch, each step reading its own
   snapshot frozen for every stage (integrate.py:323-339); the window's last
   launch also runs the FP64 FTLE epilogue (large single-launch windows process their
   seed groups in Morton order of their cells, which leaves results unchanged) with
   ``map_period = (num_steps - 1) * |dt|`` (ftle_solver.py:156,160);
4. progress messages ``(id, i)`` and ``(id, "done")`` go to the queue
   (ftle_solver.py:104-110); the field is written or returned (112-120).

With an ``AnalyticalInterpolator`` the window steps as the reference does
(ftle_solver.py:99-102 with interpolate.py:788,812): ``num_steps`` steps, each
calling the user's ``velocity_fn(0.0, points)`` on the device stage positions
at every stage, the stages combined by HIP kernels; then the FP64 FTLE kernel.
With an opt-in grid (``set_grid``) the field sampled once at time 0 is read by
the fused window kernel in steady mode instead, and the window raises if a
particle left the sampled box.
"""

from __future__ import annotations

from typing import Optional

import torch

from . import window as W
from .field import to_device, to_host
from .integrate import Integrator
from .interpolate import AnalyticalInterpolator, HighPerfInterpolator


class FTLESolver:
    def __init__(self, source, integrator: Integrator, progress_queue=None, output_writer=None,
                 launch: Optional[W.LaunchConfig] = None):
        self.source = source
        self.integrator = integrator
        self.output_writer = output_writer
        self.progress_queue = progress_queue
        self.launch = launch

    def _method(self) -> str:
        return {0: "euler", 1: "ab2", 2: "rk4"}[self.integrator.method]

    def run(self):
        self.particles = self.source.get_particles()
        id_ = self.source.id
        timestep = self.source.timestep
        num_steps = self.source.num_steps
        interp = self.integrator.interpolator
        pos = self.particles.positions
        on_host = not isinstance(pos, torch.Tensor)
        dev = to_device(pos, torch.float64 if on_host else None)
        if dev.dtype not in (torch.float32, torch.float64):
            dev = dev.to(torch.float64)
        init = dev.clone()
        map_period = (num_steps - 1) * abs(timestep)
        method = self._method()

        if isinstance(interp, AnalyticalInterpolator) and not interp.is_gridded:
            dev = dev.contiguous()
            for i in range(num_steps):
                self.source.update_interpolator(interp, i)
                self.integrator.step_device(timestep, dev)
                if self.progress_queue:
                    self.progress_queue.put((id_, i))
            ftle = W.ftle_from_groups(dev, init, map_period)
        elif isinstance(interp, AnalyticalInterpolator):
            field, k = interp.gridded().device_field()
            ftle = W.advect_ftle(field, init, method=method, time_mode="steady", n_steps=num_steps,
                                 h=timestep, map_period=map_period, first=k, positions_out=dev,
                                 cfg=self.launch)
            interp.check_inside_grid(dev)
            for i in range(num_steps):
                self.source.update_interpolator(interp, i)
                if self.progress_queue:
                    self.progress_queue.put((id_, i))
        elif isinstance(interp, HighPerfInterpolator):
            ring = interp.begin_window(num_steps)
            start = 0
            ftle = None
            vprev = None
            while start < num_steps:
                stop = min(num_steps, start + ring)
                for i in range(start, stop):
                    self.source.update_interpolator(interp, i)
                field = interp.field
                last = stop == num_steps
                if start == 0 and last:
                    n_groups = init.shape[0] // (2 * init.shape[1])
                    ftle = W.advect_ftle(field, init, method=method, time_mode="frozen",
                                         n_steps=num_steps, h=timestep, map_period=map_period,
                                         first=0, positions_out=dev, cfg=self.launch,
                                         order="morton" if n_groups >= 1 << 18 else None)
                else:
                    if method == "ab2" and vprev is None:
                        vprev = torch.zeros_like(dev)
                    W.advect(field, dev, method=method, time_mode="frozen", n_steps=stop - start,
                             h=timestep, first=start, vprev=vprev, have_prev=start > 0,
                             cfg=self.launch)
                    if last:
                        ftle = W.ftle_from_groups(dev, init, map_period)
                if self.progress_queue:
                    for i in range(start, stop):
                        self.progress_queue.put((id_, i))
                start = stop
            if ftle is None:  # zero-step window
                ftle = W.ftle_from_groups(dev, init, map_period)
        else:
            raise TypeError("FTLESolver needs a 'grid' or 'analytical' interpolator")

        if self.progress_queue:
            self.progress_queue.put((id_, "done"))

        if on_host:
            to_host(dev, out=pos)
        else:
            pos.copy_(dev)
        ftle_field = to_host(ftle) if on_host else ftle

        if self.output_writer is not None:
            self.output_writer.write(f"ftle_{id_}", ftle_field, self.particles.initial_centroid)
            return None
        return ftle_field
