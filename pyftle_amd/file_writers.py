"""FTLE field writers (file_writers.py:12-302): same factory, names and ``.mat`` keys.

``MatWriter`` saves ``ftle``, ``x``, ``y`` (, ``z``) with ``scipy.io.savemat``,
reshaped to ``(nx, ny, nz)`` (and the centroid to ``(nx, ny, nz, dim)``) when a
grid shape is given, flat otherwise (file_writers.py:92-164).  ``VTKWriter``
writes the same point arrays as the reference's pyvista-based writer
(file_writers.py:167-266) as VTK XML files of its own, so it needs no pyvista.
"""

from __future__ import annotations

import base64
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


def _vtk_data_array(a: np.ndarray, name: str | None = None, components: int = 1) -> str:
    """One ``<DataArray format="binary">``: base64 of a UInt64 byte count and the raw
    little-endian values (VTK XML inline binary encoding, ``header_type="UInt64"``)."""
    a = np.ascontiguousarray(a)
    types = {np.dtype("<f8"): "Float64", np.dtype("<f4"): "Float32", np.dtype("<i8"): "Int64",
             np.dtype("<i4"): "Int32"}
    a = a.astype(a.dtype.newbyteorder("<"), copy=False)
    vtype = types[a.dtype]
    raw = a.tobytes()
    payload = base64.b64encode(np.uint64(len(raw)).astype("<u8").tobytes() + raw).decode("ascii")
    nm = f' Name="{name}"' if name else ""
    return (f'<DataArray type="{vtype}"{nm} NumberOfComponents="{components}" format="binary">'
            f"{payload}</DataArray>")


def _vtk_file(kind: str, body: str) -> str:
    return ('<?xml version="1.0"?>\n'
            f'<VTKFile type="{kind}" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
            f"{body}\n</VTKFile>\n")


def write_vts(path: str, x: np.ndarray, y: np.ndarray, z: np.ndarray, point_data: dict) -> None:
    """StructuredGrid of dimensions ``x.shape``: point ``i + nx*(j + ny*k)`` is
    ``(x, y, z)[i, j, k]`` (Fortran order), as ``pyvista.StructuredGrid(x, y, z)`` builds it."""
    dims = list(x.shape) + [1] * (3 - x.ndim)
    pts = np.stack([np.asarray(c).ravel(order="F") for c in (x, y, z)], axis=1)
    ext = " ".join(f"0 {d - 1}" for d in dims)
    pd = "".join(_vtk_data_array(np.asarray(v), k) for k, v in point_data.items())
    first = next(iter(point_data), None)
    scal = f' Scalars="{first}"' if first else ""
    body = (f'<StructuredGrid WholeExtent="{ext}"><Piece Extent="{ext}">'
            f"<PointData{scal}>{pd}</PointData>"
            f"<Points>{_vtk_data_array(pts, components=3)}</Points>"
            "</Piece></StructuredGrid>")
    with open(path, "w") as f:
        f.write(_vtk_file("StructuredGrid", body))


def write_vtp(path: str, points: np.ndarray, point_data: dict) -> None:
    """PolyData of ``points (n, 3)`` with one vertex cell per point, as
    ``pyvista.PolyData(points)`` builds it."""
    n = points.shape[0]
    pd = "".join(_vtk_data_array(np.asarray(v), k) for k, v in point_data.items())
    first = next(iter(point_data), None)
    scal = f' Scalars="{first}"' if first else ""
    body = (f'<PolyData><Piece NumberOfPoints="{n}" NumberOfVerts="{n}" NumberOfLines="0" '
            'NumberOfStrips="0" NumberOfPolys="0">'
            f"<PointData{scal}>{pd}</PointData>"
            f"<Points>{_vtk_data_array(points, components=3)}</Points>"
            "<Verts>" + _vtk_data_array(np.arange(n, dtype=np.int64), "connectivity")
            + _vtk_data_array(np.arange(1, n + 1, dtype=np.int64), "offsets") + "</Verts>"
            "</Piece></PolyData>")
    with open(path, "w") as f:
        f.write(_vtk_file("PolyData", body))


class VTKWriter(FTLEWriter):
    """``.vts`` (structured, with ``grid_shape``) or ``.vtp`` (point cloud) files with the
    point arrays of file_writers.py:167-266, written as VTK XML directly (no pyvista):

    * 2D ``(nx, ny)``: centroids reshaped ``(nx, ny, 1, 2)`` in Fortran order, ``z = 0``,
      ``ftle`` in seed order (file_writers.py:223-236);
    * 3D ``(nx, ny, nz)``: centroids reshaped ``(nx, ny, nz, 3)`` in C order, ``ftle``
      reshaped the same way, its first two axes swapped and raveled in Fortran order
      (file_writers.py:238-252);
    * no grid shape: the centroids (``z = 0`` in 2D) with ``ftle`` as point data.
    """

    def write(self, filename: str, ftle_field, particles_centroid) -> None:
        ftle = _host(ftle_field)
        cen = _host(particles_centroid)
        if self.dim is None:
            self.dim = cen.shape[1]
        out = os.path.join(self.path, filename)
        if self.grid_shape is not None:
            if len(self.grid_shape) == 2:
                nx, ny = self.grid_shape
                c = cen.reshape(nx, ny, 1, self.dim, order="F")
                x, y = c[..., 0], c[..., 1]
                write_vts(out + ".vts", x, y, np.zeros_like(x), {"ftle": ftle.ravel(order="F")})
            elif len(self.grid_shape) == 3:
                nx, ny, nz = self.grid_shape
                c = cen.reshape(nx, ny, nz, self.dim)
                f = np.transpose(ftle.reshape((nx, ny, nz)), axes=(1, 0, 2)).ravel(order="F")
                write_vts(out + ".vts", c[..., 0], c[..., 1], c[..., 2], {"ftle": f})
            else:
                raise ValueError(f"Invalid grid_shape length {len(self.grid_shape)}. Must be 2 or 3.")
        else:
            pts = np.hstack([cen, np.zeros((cen.shape[0], 1))]) if self.dim == 2 else cen
            write_vtp(out + ".vtp", np.asarray(pts), {"ftle": ftle.ravel()})


def create_writer(output_format: str, directory_path: str, grid_shape: Optional[tuple] = None) -> FTLEWriter:
    if output_format == "mat":
        return MatWriter(directory_path, grid_shape)
    if output_format == "vtk":
        return VTKWriter(directory_path, grid_shape)
    raise ValueError(f"Unsupported output format: {output_format}")
