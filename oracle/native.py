"""Loaders for the oracle's native pieces (test infrastructure only).

* ``cpu_path()``: ctypes handle to ``oracle/cpu_path.so`` (C restatement of
  the reference CPU path, see ``cpu_path.c``).
* ``reference_ginterp()``: the reference's own ``GridInterp.cpp`` compiled into
  ``oracle/_ref`` by ``oracle/Makefile``; ``None`` when it was not built.
"""

from __future__ import annotations

import ctypes
import glob
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_cpu = None


def cpu_path():
    global _cpu
    if _cpu is None:
        path = os.path.join(HERE, "cpu_path.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/cpu_path.so is not built (run: make -C oracle)")
        lib = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.cpu_interp2d.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int64, dp]
        lib.cpu_interp3d.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int64, dp]
        lib.cpu_advect.argtypes = [
            ctypes.c_int, ip, dp, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
            ctypes.c_int, dp, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_int,
            ctypes.c_int, ctypes.c_int,
        ]
        lib.cpu_advect.restype = ctypes.c_int
        lib.cpu_num_threads.restype = ctypes.c_int
        _cpu = lib
    return _cpu


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def cpu_interp2d(v, points):
    v = np.ascontiguousarray(v, dtype=np.float64)
    points = np.ascontiguousarray(points, dtype=np.float64)
    out = np.empty(points.shape[0])
    cpu_path().cpu_interp2d(_dp(v), v.shape[0], v.shape[1], _dp(points), points.shape[0], _dp(out))
    return out


def cpu_interp3d(v, points):
    v = np.ascontiguousarray(v, dtype=np.float64)
    points = np.ascontiguousarray(points, dtype=np.float64)
    out = np.empty(points.shape[0])
    cpu_path().cpu_interp3d(_dp(v), *v.shape, _dp(points), points.shape[0], _dp(out))
    return out


METHODS = {"euler": 0, "ab2": 1, "rk4": 2}


def cpu_advect(fields, positions, t0, h, n_steps, method="rk4", mode=0, snap_t0=0.0,
               snap_dt=1.0, threads=0):
    """Advect ``positions`` (copy) through ``fields`` (list of ``GridField``).

    ``mode`` 0: step i uses snapshot i frozen (reference); 1: time-linear.
    """
    f0 = fields[0]
    dim = f0.dim
    comps = np.stack([np.stack([np.ascontiguousarray(c) for c in f.v]) for f in fields])
    comps = np.ascontiguousarray(comps.reshape(len(fields), dim, -1))
    x = np.array(positions, dtype=np.float64, copy=True, order="C")
    n = (ctypes.c_int * 3)(*[int(s) for s in f0.v[0].shape] + [1] * (3 - dim))
    mins = np.array(f0.mins + [0.0] * (3 - dim), dtype=np.float64)
    deltas = np.array(f0.deltas + [1.0] * (3 - dim), dtype=np.float64)
    rc = cpu_path().cpu_advect(
        dim, n, _dp(mins), _dp(deltas), _dp(comps), len(fields), snap_t0, snap_dt, mode,
        _dp(x), x.shape[0], t0, h, n_steps, METHODS[method], threads,
    )
    if rc != 0:
        raise MemoryError("cpu_advect allocation failed")
    return x


_ref_mod = None


def reference_ginterp():
    """Import the compiled reference ``ginterp`` module from ``oracle/_ref``."""
    global _ref_mod
    if _ref_mod is None:
        hits = glob.glob(os.path.join(HERE, "_ref", "ginterp*.so"))
        if not hits:
            return None
        spec = importlib.util.spec_from_file_location("ginterp", hits[0])
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _ref_mod = mod
    return _ref_mod
