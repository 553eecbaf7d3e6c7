"""MI355X (gfx950) implementation of the RRGalaxSim solver's per-step hot path.

The compute lives in ``librrg.so`` (hand-written HIP kernels behind the C ABI
declared in ``include/rrg.h``); ``rrgsim`` is the params.json-driven CLI.
This package is the Python view of that ABI (ctypes), plus seeded initial
conditions and the ``torch.autograd.Function`` binding of the adjoint.
"""
from . import _lib
from ._lib import RRGError, lib
from .solver import (Simulation, cic_deposit, default_params, direct_accel, direct_grav, gather, gradient,
                     make_params, ngp_bin, parse_params_json, read_ini, steady_err, wave_iterate, wave_schedule)

__all__ = ["RRGError", "Simulation", "cic_deposit", "default_params", "direct_accel", "direct_grav", "gather",
           "gradient", "lib", "make_params", "ngp_bin", "parse_params_json", "read_ini", "steady_err", "wave_iterate",
           "wave_schedule",
           "_lib"]
