"""Integrator strategies (integrate.py:117-387) on the fused HIP kernel.

``integrate(h, particles)`` advances ``particles.positions`` by one step in
place, exactly like the reference strategies, by launching the advection
kernel with ``n_steps=1`` on the interpolator's current snapshot.  The window
driver (``FTLESolver``) instead launches the same kernel once for all steps of
a window (``advect``).

AB2 keeps its previous velocity on the integrator (integrate.py:220,245-257):
the first call is an Euler step, later calls use ``1.5 v - 0.5 v_prev``.

With an ``AnalyticalInterpolator`` (and no opt-in grid) a step is taken the
reference's way: the user's ``velocity_fn`` is evaluated on the device stage
positions at every stage (k1..k4 for RK4), and ``pyftle_stage_combine`` forms
the stage positions and the update in the reference's operation order
(integrate.py:16-109, 291-342).
"""

from __future__ import annotations

import ctypes
from abc import ABC, abstractmethod

import numpy as np
import torch

from . import _native as N
from .field import dtype_code, ptr, stream_handle, to_device
from .interpolate import AnalyticalInterpolator, HighPerfInterpolator, Interpolator


def _opts(method, time_mode, n_steps, first=0, t_start=0.0, h=0.0, have_prev=False, lane_order=0,
          groups_per_block=0, xcd_remap=1) -> N.AdvectOpts:
    o = N.AdvectOpts()
    o.method = method
    o.time_mode = time_mode
    o.n_steps = int(n_steps)
    o.first_snapshot = int(first)
    o.t_start = float(t_start)
    o.h = float(h)
    o.ab2_have_prev = int(bool(have_prev))
    o.lane_order = int(lane_order)
    o.groups_per_block = int(groups_per_block)
    o.xcd_remap = int(xcd_remap)
    return o


class Integrator(ABC):
    """integrate.py:117-153."""

    method = N.RK4

    def __init__(self, interpolator: Interpolator, **kwargs) -> None:
        self.interpolator = interpolator

    def _grid(self) -> HighPerfInterpolator:
        interp = self.interpolator
        if isinstance(interp, AnalyticalInterpolator):
            return interp.gridded()
        if not isinstance(interp, HighPerfInterpolator):
            raise TypeError("device integrators need a 'grid' or 'analytical' interpolator")
        return interp

    def stage_evaluated(self) -> bool:
        """True when steps call the analytical velocity function at every stage."""
        interp = self.interpolator
        return isinstance(interp, AnalyticalInterpolator) and not interp.is_gridded

    def _combine(self, op: int, dev: torch.Tensor, c: float, k1, k2=None, k3=None, k4=None, out=None):
        N.call("pyftle_stage_combine", op, dtype_code(dev.dtype), dev.numel(), float(c), ptr(dev),
               ptr(k1), ptr(k2) if k2 is not None else None, ptr(k3) if k3 is not None else None,
               ptr(k4) if k4 is not None else None, ptr(out) if out is not None else None, stream_handle())

    def step_device(self, h: float, dev: torch.Tensor) -> None:
        """One step of the analytical strategy on a contiguous device tensor, in place."""
        raise NotImplementedError

    def _device_positions(self, particles):
        pos = particles.positions
        on_host = not isinstance(pos, torch.Tensor)
        dev = to_device(pos, torch.float64 if on_host else None)
        if dev.dtype not in (torch.float32, torch.float64):
            dev = dev.to(torch.float64)
        return pos, on_host, dev.contiguous()

    def _write_back(self, pos, on_host, dev) -> None:
        if on_host:
            np.copyto(pos, dev.cpu().numpy())
        elif dev.data_ptr() != pos.data_ptr():
            pos.copy_(dev)

    def _step(self, h: float, particles, vprev=None, have_prev=False) -> None:
        if self.stage_evaluated():
            pos, on_host, dev = self._device_positions(particles)
            self.step_device(h, dev)
            self._write_back(pos, on_host, dev)
            return
        grid = self._grid()
        field, k = grid.device_field()
        pos = particles.positions
        on_host = not isinstance(pos, torch.Tensor)
        dev = to_device(pos, torch.float64 if on_host else None)
        if dev.dtype not in (torch.float32, torch.float64):
            dev = dev.to(torch.float64)
        opts = _opts(self.method, N.TIME_FROZEN, 1, first=k, h=h, have_prev=have_prev)
        vp = ptr(vprev) if vprev is not None else ctypes.c_void_p(0)
        N.call("pyftle_advect", ctypes.byref(field.desc), ctypes.byref(opts), ptr(dev),
               dtype_code(dev.dtype), dev.shape[0], vp, stream_handle())
        if on_host:
            np.copyto(pos, dev.cpu().numpy())
        elif dev.data_ptr() != pos.data_ptr():
            pos.copy_(dev)

    @abstractmethod
    def integrate(self, h: float, particles) -> None:
        pass


class EulerIntegrator(Integrator):
    """integrate.py:161-196: ``x += h v``."""

    method = N.EULER

    def integrate(self, h: float, particles) -> None:
        self._step(h, particles)

    def step_device(self, h: float, dev: torch.Tensor) -> None:
        v = self.interpolator.evaluate(dev)
        self._combine(N.COMBINE_EULER, dev, h, v)


class AdamsBashforth2Integrator(Integrator):
    """integrate.py:204-257: Euler first, then ``x += h (1.5 v - 0.5 v_prev)``."""

    method = N.AB2

    def __init__(self, interpolator: Interpolator, **kwargs) -> None:
        super().__init__(interpolator, **kwargs)
        self.previous_velocity = None  # device tensor, same shape as positions

    def integrate(self, h: float, particles) -> None:
        if self.stage_evaluated():
            self._step(h, particles)
            return
        pos = particles.positions
        dt = pos.dtype if isinstance(pos, torch.Tensor) and pos.dtype in (torch.float32, torch.float64) \
            else torch.float64
        have_prev = self.previous_velocity is not None and tuple(self.previous_velocity.shape) == tuple(pos.shape)
        if not have_prev:
            self.previous_velocity = torch.zeros(tuple(pos.shape), dtype=dt, device=torch.device("cuda"))
        self._step(h, particles, vprev=self.previous_velocity, have_prev=have_prev)

    def step_device(self, h: float, dev: torch.Tensor) -> None:
        v = self.interpolator.evaluate(dev)
        prev = self.previous_velocity
        if prev is None or tuple(prev.shape) != tuple(dev.shape) or prev.dtype != dev.dtype:
            # first call: Euler, then keep v (integrate.py:248-257)
            self.previous_velocity = torch.empty_like(dev)
            self._combine(N.COMBINE_EULER, dev, h, v, out=self.previous_velocity)
        else:
            self._combine(N.COMBINE_AB2, dev, h, v, prev, out=prev)


class RungeKutta4Integrator(Integrator):
    """integrate.py:265-342: classic RK4, field frozen across the stages."""

    method = N.RK4

    def __init__(self, interpolator: Interpolator, **kwargs) -> None:
        super().__init__(interpolator, **kwargs)
        self._tmp = None  # stage-position buffer of the analytical path (integrate.py:316-321)

    def integrate(self, h: float, particles) -> None:
        self._step(h, particles)

    def step_device(self, h: float, dev: torch.Tensor) -> None:
        f = self.interpolator.evaluate
        if self._tmp is None or self._tmp.shape != dev.shape or self._tmp.dtype != dev.dtype:
            self._tmp = torch.empty_like(dev)
        tmp = self._tmp
        k1 = f(dev)
        self._combine(N.COMBINE_STAGE, dev, 0.5 * h, k1, out=tmp)
        k2 = f(tmp)
        self._combine(N.COMBINE_STAGE, dev, 0.5 * h, k2, out=tmp)
        k3 = f(tmp)
        self._combine(N.COMBINE_STAGE, dev, h, k3, out=tmp)
        k4 = f(tmp)
        self._combine(N.COMBINE_RK4, dev, h / 6.0, k1, k2, k3, k4)


INTEGRATORS = {
    "ab2": AdamsBashforth2Integrator,
    "euler": EulerIntegrator,
    "rk4": RungeKutta4Integrator,
}


def create_integrator(integrator_name: str, interpolator: Interpolator) -> Integrator:
    """integrate.py:350-387 (case-insensitive; same ``ValueError`` text)."""
    integrator_name = integrator_name.lower()
    if integrator_name not in INTEGRATORS:
        raise ValueError(
            f"Invalid integrator name '{integrator_name}'. "
            f"Choose from {list(INTEGRATORS.keys())}."
        )
    return INTEGRATORS[integrator_name](interpolator)


def method_code(name_or_integrator) -> int:
    if isinstance(name_or_integrator, Integrator):
        return name_or_integrator.method
    return {"euler": N.EULER, "ab2": N.AB2, "rk4": N.RK4}[str(name_or_integrator).lower()]
