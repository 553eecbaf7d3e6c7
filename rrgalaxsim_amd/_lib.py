"""ctypes binding of librrg.so (include/rrg.h).

The product's compute path is the HIP library; this module only moves host
buffers across the C ABI.  It raises if the library is missing -- there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librrg.so")
CLI_PATH = os.path.join(_HERE, "rrgsim")

RRG_OK = 0
RRG_ERR_INVALID = -1
RRG_ERR_HIP = -2
RRG_ERR_IO = -3
RRG_ERR_STATE = -4

REAL_F64 = 0
REAL_F32 = 1
FORCE_F64 = 0
FORCE_MIXED = 1
FORCE_FP32 = 2
FORCE_F64_EXACT = 3
FORCE_KINDS = {"fp64": 0, "mixed": 1, "fp32": 2, "fp64_exact": 3}
BC_ZERO_DIRICHLET = 0
BC_PERIODIC = 1
BC_MONOPOLE_DIRICHLET = 2
BC_MONOPOLE_RADIATIVE = 3
BOUNDARIES = {"zero_dirichlet": BC_ZERO_DIRICHLET, "periodic": BC_PERIODIC, "monopole_dirichlet": BC_MONOPOLE_DIRICHLET,
              "monopole_radiative": BC_MONOPOLE_RADIATIVE}

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_vp = ctypes.c_void_p


class SimParams(ctypes.Structure):
    _fields_ = [("time_max", _f64), ("dt_save", _f64), ("dt_dynamics", _f64),
                ("saves_without_dump", ctypes.c_uint64)]


class GridParams(ctypes.Structure):
    _fields_ = [("nx", _i32), ("dx", _f64), ("bc_frac", _f64)]


class WaveParams(ctypes.Structure):
    _fields_ = [("diss_base", _f64), ("diss_extra", _f64), ("wave_speed", _f64), ("dt", _f64),
                ("setup_iterations", _i32), ("iterations", _i32), ("tolerance", _f64),
                ("print_step", _i32), ("steady_check_step", _i32), ("steady_hold_steps", _i32)]


class IniParams(ctypes.Structure):
    _fields_ = [("mass_star", _f64), ("mass_dark", _f64), ("soft_star", _f64), ("soft_dark", _f64),
                ("ini_file_star", ctypes.c_char * 1024), ("ini_file_dark", ctypes.c_char * 1024)]


class Params(ctypes.Structure):
    _fields_ = [("sim", SimParams), ("has_grid", _i32), ("grid", GridParams), ("has_wave", _i32),
                ("wave", WaveParams), ("use_wave_model", _i32), ("real_type", _i32), ("force_kind", _i32),
                ("boundary", _i32), ("fixed_exponent", _i32), ("device", _i32), ("use_graphs", _i32),
                ("autotune", _i32), ("wave_fused", _i32), ("sort_interval", _i32)]


class Conservation(ctypes.Structure):
    _fields_ = [("time", _f64), ("Ek", _f64), ("Ep", _f64), ("momentum", _f64 * 3), ("angular", _f64 * 3)]


class Stats(ctypes.Structure):
    _fields_ = [("n", _i64), ("steps", _i64), ("last_wave_iterations", _i32), ("fixed_exponent", _i32),
                ("time", _f64), ("systems", _i32), ("tile_ib", _i32), ("slice_len", _i64), ("workgroups", _i32),
                ("work_items", _i32), ("tuned", _i32), ("tune_candidates", _i32), ("graph_launches", _i64)]


class Timing(ctypes.Structure):
    _fields_ = [("ms", _f64 * 12), ("launches", _i64 * 12)]


STAGES = ["deposit", "source", "wave", "gradient", "gather", "direct", "predict", "correct", "sort", "adj_force",
          "adj_wave", "adj_particle"]


# Every symbol the header declares, with its ctypes signature.
SIGNATURES = {
    "rrg_last_error": (ctypes.c_char_p, []),
    "rrg_version": (ctypes.c_char_p, []),
    "rrg_device_count": (ctypes.c_int, []),
    "rrg_params_default": (ctypes.c_int, [ctypes.POINTER(Params)]),
    "rrg_params_finalize": (ctypes.c_int, [ctypes.POINTER(Params)]),
    "rrg_params_parse_json": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(Params), ctypes.POINTER(IniParams)]),
    "rrg_ini_read": (ctypes.c_int, [ctypes.POINTER(IniParams), ctypes.c_char_p, ctypes.POINTER(_i64),
                                    _vp, _vp, _vp, _vp]),
    "rrg_create": (ctypes.c_int, [ctypes.POINTER(Params), _i64, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
    "rrg_create_ensemble": (ctypes.c_int, [ctypes.POINTER(Params), _i32, _vp, _vp, _vp, _vp, _vp,
                                           ctypes.POINTER(_vp)]),
    "rrg_destroy": (ctypes.c_int, [_vp]),
    "rrg_get_conservation_systems": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(Conservation)]),
    "rrg_adjoint_begin": (ctypes.c_int, [_vp, _i32]),
    "rrg_adjoint": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "rrg_precalc": (ctypes.c_int, [_vp]),
    "rrg_step": (ctypes.c_int, [_vp, _i32]),
    "rrg_sync": (ctypes.c_int, [_vp]),
    "rrg_get_state": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "rrg_set_state": (ctypes.c_int, [_vp, _vp, _vp]),
    "rrg_get_state_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "rrg_wave_schedule": (ctypes.c_int, [_i32, _i32, _i32, _i32, _i32, _vp]),
    "rrg_set_state_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "rrg_adjoint_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "rrg_get_conservation": (ctypes.c_int, [_vp, ctypes.POINTER(Conservation)]),
    "rrg_get_grid": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "rrg_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(Stats)]),
    "rrg_synchronize": (ctypes.c_int, [_vp]),
    "rrg_set_timing": (ctypes.c_int, [_vp, _i32]),
    "rrg_get_timing": (ctypes.c_int, [_vp, ctypes.POINTER(Timing)]),
    "rrg_op_direct_accel": (ctypes.c_int, [_i32, _i64, _vp, _vp, _vp, _vp]),
    "rrg_op_direct_grav": (ctypes.c_int, [_i32, _i64, _vp, _vp, _vp, _vp]),
    "rrg_op_cic_deposit": (ctypes.c_int, [_i32, _i64, _vp, _vp, _i32, _f64, _f64, _i32, _vp]),
    "rrg_op_wave_iterate": (ctypes.c_int, [_i32, _i32, _f64, _f64, _f64, _f64, _i32, _f64, _i32, _vp, _vp, _vp]),
    "rrg_op_gradient": (ctypes.c_int, [_i32, _i32, _f64, _vp, _vp]),
    "rrg_op_gather": (ctypes.c_int, [_i32, _i64, _vp, _i32, _f64, _f64, _vp, _vp]),
    "rrg_op_ngp_bin": (ctypes.c_int, [_i32, _i64, _vp, _i32, _f64, _f64, _vp, _vp, _vp, _vp]),
    "rrg_op_steady_err": (ctypes.c_int, [_i32, _i64, _vp, _vp, ctypes.POINTER(_f64)]),
    "rrg_op_direct_accel_noalloc": (ctypes.c_int, [_i32, _i64, _vp, _vp, _vp, _vp, _vp]),
}

_lib = None


class RRGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Load librrg.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                               " or `make -C rrgalaxsim_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != RRG_OK:
        raise RRGError(rc, lib().rrg_last_error().decode())


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_vp)


def f64(a, shape=None) -> np.ndarray:
    out = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        out = out.reshape(shape)
    return out
