"""Cauchy-Green tensor, largest eigenvalue and FTLE (ftle.py:9-233), FP64 on device.

* ``compute_cauchy_green_2x2``: ``C = F^T F`` (ftle.py:62-67);
  ``compute_cauchy_green_3x3``: ``C = F F^T`` (ftle.py:96).
* ``max_eigenvalue_2x2``: ``0.5*(tr + sqrt((A-D)^2 + 4 B C))``, algebraically
  the reference's ``0.5*(tr + sqrt(tr^2 - 4 det))`` (ftle.py:155-161) but
  free of the cancellation that turns a near-isotropic C into NaN there.
* ``max_eigenvalue_3x3``: symmetric trigonometric closed form in place of
  ``np.linalg.eigvals`` (ftle.py:118-119); C from ``compute_cauchy_green_3x3``
  is symmetric.
* ``compute_ftle_*``: ``(1/T) * log(sqrt(lambda_max))`` (ftle.py:198,233).

NumPy in -> NumPy out; tensor in -> tensor out.
"""

from __future__ import annotations

import torch

from . import _native as N
from .field import ptr, stream_handle, to_device


def _dev(a):
    is_np = not isinstance(a, torch.Tensor)
    t = to_device(a, torch.float64)
    return t, is_np


def _ret(t, is_np):
    return t.cpu().numpy() if is_np else t


def _cg(F, dim):
    t, is_np = _dev(F)
    if t.dim() != 3 or t.shape[1:] != (dim, dim):
        raise ValueError(f"flow_map_jacobian must have shape (N, {dim}, {dim})")
    out = torch.empty_like(t)
    N.call("pyftle_cauchy_green", ptr(t), dim, t.shape[0], ptr(out), stream_handle())
    return _ret(out, is_np)


def _lam(C, dim):
    t, is_np = _dev(C)
    if t.dim() != 3 or t.shape[1:] != (dim, dim):
        raise ValueError(f"cauchy_green_tensor must have shape (N, {dim}, {dim})")
    out = torch.empty(t.shape[0], dtype=torch.float64, device=t.device)
    N.call("pyftle_max_eigenvalue", ptr(t), dim, t.shape[0], ptr(out), stream_handle())
    return _ret(out, is_np)


def _ftle(F, T, dim):
    t, is_np = _dev(F)
    if t.dim() != 3 or t.shape[1:] != (dim, dim):
        raise ValueError(f"flow_map_jacobian must have shape (N, {dim}, {dim})")
    out = torch.empty(t.shape[0], dtype=torch.float64, device=t.device)
    N.call("pyftle_ftle_from_jacobian", ptr(t), dim, t.shape[0], float(T), ptr(out), stream_handle())
    return _ret(out, is_np)


def compute_cauchy_green_2x2(flow_map_jacobian):
    return _cg(flow_map_jacobian, 2)


def compute_cauchy_green_3x3(flow_map_jacobian):
    return _cg(flow_map_jacobian, 3)


def max_eigenvalue_2x2(cauchy_green_tensor):
    return _lam(cauchy_green_tensor, 2)


def max_eigenvalue_3x3(cauchy_green_tensor):
    return _lam(cauchy_green_tensor, 3)


def compute_ftle_2x2(flow_map_jacobian, map_period: float):
    return _ftle(flow_map_jacobian, map_period, 2)


def compute_ftle_3x3(flow_map_jacobian, map_period: float):
    return _ftle(flow_map_jacobian, map_period, 3)
