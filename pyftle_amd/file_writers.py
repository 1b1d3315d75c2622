"""FTLE field writers (file_writers.py:12-302): same factory, names and ``.mat`` keys.

``MatWriter`` saves ``ftle``, ``x``, ``y`` (, ``z``) with ``scipy.io.savemat``,
reshaped to ``(nx, ny, nz)`` (and the centroid to ``(nx, ny, nz, dim)``) when a
grid shape is given, flat otherwise (file_writers.py:92-164).  ``VTKWriter``
needs pyvista, as the reference's does (file_writers.py:6,167-266); without it
the writer raises at construction.
"""

from __future__ import annotations

import os
from abc import ABC, abstractmethod
from typing import Optional

import numpy as np
from scipy.io import savemat


def _host(a) -> np.ndarray:
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


class FTLEWriter(ABC):
    def __init__(self, directory_path, grid_shape: Optional[tuple] = None) -> None:
        self.path = directory_path
        try:
            os.makedirs(self.path, exist_ok=True)
        except OSError as e:
            print(f"Error creating output folder: {e}")
        self.grid_shape = grid_shape
        self.dim: Optional[int] = None

    @abstractmethod
    def write(self, filename: str, ftle_field, particles_centroid) -> None: ...


class MatWriter(FTLEWriter):
    def write(self, filename: str, ftle_field, particles_centroid) -> None:
        ftle = _host(ftle_field)
        cen = _host(particles_centroid)
        if self.dim is None:
            self.dim = cen.shape[1]
        out = os.path.join(self.path, filename + ".mat")
        if self.grid_shape:
            if len(self.grid_shape) == 2:
                nx, ny = self.grid_shape
                nz = 1
            elif len(self.grid_shape) == 3:
                nx, ny, nz = self.grid_shape
            else:
                raise ValueError(f"Invalid grid_shape length {len(self.grid_shape)}. Must be 2 or 3.")
            ftle = ftle.reshape(nx, ny, nz)
            cen = cen.reshape(nx, ny, nz, self.dim)
            data = {"ftle": ftle, "x": cen[..., 0], "y": cen[..., 1]}
            if self.dim == 3:
                data["z"] = cen[..., 2]
        else:
            data = {"ftle": ftle.ravel(), "x": cen[:, 0], "y": cen[:, 1]}
            if self.dim == 3:
                data["z"] = cen[:, 2]
        savemat(out, data)


class VTKWriter(FTLEWriter):
    def __init__(self, directory_path, grid_shape: Optional[tuple] = None) -> None:
        try:
            import pyvista  # noqa: F401
        except ImportError as e:
            raise RuntimeError("VTK output needs pyvista, which is not installed; use output_format='mat'") from e
        super().__init__(directory_path, grid_shape)

    def write(self, filename: str, ftle_field, particles_centroid) -> None:
        import pyvista as pv

        ftle = _host(ftle_field)
        cen = _host(particles_centroid)
        if self.dim is None:
            self.dim = cen.shape[1]
        out = os.path.join(self.path, filename)
        if self.grid_shape:
            shape = tuple(self.grid_shape) + ((1,) if len(self.grid_shape) == 2 else ())
            c = cen.reshape(*shape, self.dim)
            z = c[..., 2] if self.dim == 3 else np.zeros(shape)
            grid = pv.StructuredGrid(c[..., 0], c[..., 1], z)
            grid["ftle"] = ftle.reshape(shape).ravel(order="F")
            grid.save(out + ".vts")
        else:
            pts = cen if self.dim == 3 else np.column_stack([cen, np.zeros(len(cen))])
            mesh = pv.PolyData(pts)
            mesh["ftle"] = ftle.ravel()
            mesh.save(out + ".vtp")


def create_writer(output_format: str, directory_path: str, grid_shape: Optional[tuple] = None) -> FTLEWriter:
    if output_format == "mat":
        return MatWriter(directory_path, grid_shape)
    if output_format == "vtk":
        return VTKWriter(directory_path, grid_shape)
    raise ValueError(f"Unsupported output format: {output_format}")
