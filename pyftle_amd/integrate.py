"""Integrator strategies (integrate.py:117-387) on the fused HIP kernel.

``integrate(h, particles)`` advances ``particles.positions`` by one step in
place, exactly like the reference strategies, by launching the advection
kernel with ``n_steps=1`` on the interpolator's current snapshot.  The window
driver (``FTLESolver``) instead launches the same kernel once for all steps of
a window (``advect``).

AB2 keeps its previous velocity on the integrator (integrate.py:220,245-257):
the first call is an Euler step, later calls use ``1.5 v - 0.5 v_prev``.
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

    def _grid(self, particles) -> HighPerfInterpolator:
        interp = self.interpolator
        if isinstance(interp, AnalyticalInterpolator):
            return interp.gridded(particles.positions)
        if not isinstance(interp, HighPerfInterpolator):
            raise TypeError("device integrators need a 'grid' or 'analytical' interpolator")
        return interp

    def _step(self, h: float, particles, vprev=None, have_prev=False) -> None:
        grid = self._grid(particles)
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


class AdamsBashforth2Integrator(Integrator):
    """integrate.py:204-257: Euler first, then ``x += h (1.5 v - 0.5 v_prev)``."""

    method = N.AB2

    def __init__(self, interpolator: Interpolator, **kwargs) -> None:
        super().__init__(interpolator, **kwargs)
        self.previous_velocity = None  # device tensor, same shape as positions

    def integrate(self, h: float, particles) -> None:
        pos = particles.positions
        dt = pos.dtype if isinstance(pos, torch.Tensor) and pos.dtype in (torch.float32, torch.float64) \
            else torch.float64
        have_prev = self.previous_velocity is not None and tuple(self.previous_velocity.shape) == tuple(pos.shape)
        if not have_prev:
            self.previous_velocity = torch.zeros(tuple(pos.shape), dtype=dt, device=torch.device("cuda"))
        self._step(h, particles, vprev=self.previous_velocity, have_prev=have_prev)


class RungeKutta4Integrator(Integrator):
    """integrate.py:265-342: classic RK4, field frozen across the stages."""

    method = N.RK4

    def integrate(self, h: float, particles) -> None:
        self._step(h, particles)


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
