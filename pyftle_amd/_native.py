"""ctypes binding of ``libpyftle_hip.so`` (declared in ``include/pyftle_hip.h``).

The library is mandatory: importing a compute function without it raises
``RuntimeError`` naming the missing file, so no code path can silently fall
back to a CPU implementation.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libpyftle_hip.so")

F32, F64 = 0, 1
LAYOUT_SINGLE, LAYOUT_PAIR, LAYOUT_CELL, LAYOUT_CELL_PAIR = 0, 1, 2, 3
EULER, AB2, RK4 = 0, 1, 2
TIME_STEADY, TIME_FROZEN, TIME_LINEAR = 0, 1, 2
ORDER_SEED_MAJOR, ORDER_NEIGHBOR_MAJOR = 0, 1
COMPUTE_NATIVE, COMPUTE_MIXED, COMPUTE_F64 = 0, 1, 2
INTERP_GINTERP, INTERP_RGI_LINEAR, INTERP_RGI_NEAREST, INTERP_RGI_CUBIC = 0, 1, 2, 3
INTERP_SCATTER_LINEAR, INTERP_SCATTER_NEAREST, INTERP_SCATTER_CUBIC = 4, 5, 6
OK, EINVAL, EHIP = 0, -1, -2
COMBINE_STAGE, COMBINE_RK4, COMBINE_EULER, COMBINE_AB2 = 0, 1, 2, 3


class Field(ctypes.Structure):
    """``pyftle_field`` (include/pyftle_hip.h)."""

    _fields_ = [
        ("dim", ctypes.c_int32),
        ("n", ctypes.c_int32 * 3),
        ("origin", ctypes.c_double * 3),
        ("spacing", ctypes.c_double * 3),
        ("periodic", ctypes.c_int32 * 3),
        ("dtype", ctypes.c_int32),
        ("layout", ctypes.c_int32),
        ("ring_slots", ctypes.c_int32),
        ("n_snapshots", ctypes.c_int32),
        ("slot_stride", ctypes.c_int64),
        ("t0", ctypes.c_double),
        ("dt_snap", ctypes.c_double),
        ("data", ctypes.c_void_p),
        ("interp", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("upper", ctypes.c_double * 3),
        ("scatter", ctypes.c_void_p),
    ]


class Scatter(ctypes.Structure):
    """``pyftle_scatter`` (include/pyftle_hip.h): device arrays of a scattered-data field."""

    _fields_ = [
        ("n_points", ctypes.c_int64),
        ("points", ctypes.c_void_p),
        ("n_simplices", ctypes.c_int64),
        ("simplices", ctypes.c_void_p),
        ("transform", ctypes.c_void_p),
        ("hull", ctypes.c_void_p),
        ("simplex_dims", ctypes.c_int32 * 3),
        ("simplex_origin", ctypes.c_double * 3),
        ("simplex_size", ctypes.c_double * 3),
        ("simplex_offsets", ctypes.c_void_p),
        ("simplex_ids", ctypes.c_void_p),
        ("point_dims", ctypes.c_int32 * 3),
        ("point_origin", ctypes.c_double * 3),
        ("point_size", ctypes.c_double * 3),
        ("point_offsets", ctypes.c_void_p),
        ("point_ids", ctypes.c_void_p),
        ("ct_g", ctypes.c_void_p),
    ]


class AdvectOpts(ctypes.Structure):
    """``pyftle_advect_opts`` (include/pyftle_hip.h)."""

    _fields_ = [
        ("method", ctypes.c_int32),
        ("time_mode", ctypes.c_int32),
        ("n_steps", ctypes.c_int32),
        ("first_snapshot", ctypes.c_int32),
        ("t_start", ctypes.c_double),
        ("h", ctypes.c_double),
        ("ab2_have_prev", ctypes.c_int32),
        ("lane_order", ctypes.c_int32),
        ("groups_per_block", ctypes.c_int32),
        ("xcd_remap", ctypes.c_int32),
        ("compute", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("seed_order", ctypes.c_void_p),
    ]


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_FP = ctypes.POINTER(Field)
_OP = ctypes.POINTER(AdvectOpts)

# name -> argtypes, in header order
SIGNATURES = {
    "pyftle_version": [],
    "pyftle_last_error": [],
    "pyftle_device_sync": [],
    "pyftle_interp2d_vec": [_vp, _i32, _i32, _vp, _i64, _vp, _vp],
    "pyftle_interp3d_vec": [_vp, _i32, _i32, _i32, _vp, _i64, _vp, _vp],
    "pyftle_pack_snapshot": [_FP, _i32, _vp, _i32, ctypes.POINTER(ctypes.c_int64), _vp],
    "pyftle_spline_work_bytes": [_FP],
    "pyftle_pack_spline": [_FP, _i32, _vp, _i32, ctypes.POINTER(ctypes.c_int64), _vp, _vp],
    "pyftle_field_sample": [_FP, _i32, _i32, _f64, _vp, _i32, _i64, _vp, _vp],
    "pyftle_advect": [_FP, _OP, _vp, _i32, _i64, _vp, _vp],
    "pyftle_stage_combine": [_i32, _i32, _i64, _f64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pyftle_seed_morton_codes": [_vp, _i32, _i64, _i32, _i32, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_double), _vp, _vp],
    "pyftle_advect_ftle": [_FP, _OP, _vp, _i32, _i64, _vp, _f64, _vp, _vp],
    "pyftle_advect_ftle_count_loads": [_FP, _OP, _vp, _i32, _i64, _vp, _f64, _vp, _vp, _vp],
    "pyftle_probe_load_rate": [_vp, _i32, _i32, _vp, _vp],
    "pyftle_probe_issue_rates": [_i32, _i32, _i32, _vp, _vp, _vp],
    "pyftle_flow_map_jacobian": [_vp, _vp, _i32, _i32, _i64, _vp, _vp],
    "pyftle_cauchy_green": [_vp, _i32, _i64, _vp, _vp],
    "pyftle_max_eigenvalue": [_vp, _i32, _i64, _vp, _vp],
    "pyftle_ftle_from_jacobian": [_vp, _i32, _i64, _f64, _vp, _vp],
    "pyftle_ftle_from_groups": [_vp, _vp, _i32, _i32, _i64, _f64, _vp, _vp],
}

_lib = None


def load(path: str | None = None):
    """Load the HIP library (raises ``RuntimeError`` when it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"HIP library {p} is not built; run `python -m pyftle_amd.build` "
            "(there is no CPU fallback)."
        )
    lib = ctypes.CDLL(p)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.pyftle_last_error.restype = ctypes.c_char_p
    lib.pyftle_spline_work_bytes.restype = ctypes.c_int64
    if path is None:
        _lib = lib
    return lib


def extra_signatures(sigs: dict):
    """Register further entry points (used by optional modules)."""
    lib = load()
    for name, argtypes in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    SIGNATURES.update(sigs)
    return lib


class PyftleError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> None:
    if rc != OK:
        msg = load().pyftle_last_error().decode(errors="replace")
        raise PyftleError(f"{what}: {msg}" if what else msg)


def call(name: str, *args) -> None:
    lib = load()
    check(getattr(lib, name)(*args), name)
