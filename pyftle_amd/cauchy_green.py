"""Flow-map Jacobian (cauchy_green.py:15-210) on the device.

``J[:, r, c] = delta_c[:, r] / initial_delta_c[:, c]`` with the neighbour
pairs right-left, top-bottom, front-back (cauchy_green.py:70-74,140-151),
computed by ``pyftle_flow_map_jacobian`` in FP64.  Inputs and outputs follow
the caller's container: NumPy in -> NumPy out, tensor in -> tensor out.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as N
from .field import dtype_code, ptr, stream_handle, to_device


def _jacobian(particles, dim: int):
    pos = particles.positions
    is_np = not isinstance(pos, torch.Tensor)
    p = to_device(pos, torch.float64 if is_np else None)
    if p.dtype not in (torch.float32, torch.float64):
        p = p.to(torch.float64)
    p0 = to_device(particles.initial_positions, p.dtype)
    if p.shape[1] != dim:
        raise ValueError(f"positions must have {dim} columns")
    n = p.shape[0] // (2 * dim)
    J = torch.empty((n, dim, dim), dtype=torch.float64, device=p.device)
    N.call("pyftle_flow_map_jacobian", ptr(p), ptr(p0), dtype_code(p.dtype), dim, n, ptr(J),
           stream_handle())
    return J.cpu().numpy() if is_np else J


def compute_flow_map_jacobian_2x2(particles) -> np.ndarray:
    """cauchy_green.py:157-181."""
    return _jacobian(particles, 2)


def compute_flow_map_jacobian_3x3(particles) -> np.ndarray:
    """cauchy_green.py:184-210."""
    return _jacobian(particles, 3)
