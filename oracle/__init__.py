"""CPU oracle for the RRGalaxSim hot path -- TEST INFRASTRUCTURE ONLY.

Thin ctypes/numpy front end over ``oracle/_build/liboracle.so`` (built from
``oracle/rrg_oracle.cpp`` by ``oracle/Makefile``).  Each function names the
reference routine it restates; the C++ file carries the file:line citations.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``rrgalaxsim_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_int = ctypes.c_int
_vp = ctypes.c_void_p


class PmConfig(ctypes.Structure):
    """Mirror of the C++ ``PmConfig`` (PM run description)."""

    _fields_ = [
        ("nx", ctypes.c_int),
        ("dx", ctypes.c_double),
        ("bc_frac", ctypes.c_double),
        ("diss_base", ctypes.c_double),
        ("wave_speed", ctypes.c_double),
        ("tolerance", ctypes.c_double),
        ("setup_iterations", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("check_step", ctypes.c_int),
        ("hold", ctypes.c_int),
        ("boundary", ctypes.c_int),
        ("fixed_e", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("diss_extra", ctypes.c_double),
    ]


BOUNDARY = {"zero_dirichlet": 0, "periodic": 1, "monopole_dirichlet": 2, "monopole_radiative": 3}


def _bc(periodic, boundary):
    if boundary is None:
        return 1 if periodic else 0
    return BOUNDARY[boundary] if isinstance(boundary, str) else int(boundary)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        for suf, ct in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            getattr(_lib, f"orc_wave_dt_{suf}").restype = ct
            getattr(_lib, f"orc_wave_dt_{suf}").argtypes = [ct, ct]
            getattr(_lib, f"orc_rho_mean_{suf}").restype = ct
            getattr(_lib, f"orc_rho_mean_{suf}").argtypes = [ct, _int, ct]
            getattr(_lib, f"orc_steady_err_{suf}").restype = ctypes.c_double
            getattr(_lib, f"orc_solve_wave_{suf}").restype = _int
    return _lib


def set_threads(n: int) -> int:
    """Sets the OpenMP threads of the parallel routines; returns the previous count."""
    prev = int(lib().orc_get_threads())
    lib().orc_set_threads(_int(int(n)))
    return prev


def _suf(dtype) -> str:
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "f64"
    if dtype == np.float32:
        return "f32"
    raise TypeError(f"unsupported dtype {dtype}")


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    return a.ctypes.data_as(_vp)


def _r(x, dtype):
    return ctypes.c_double(x) if np.dtype(dtype) == np.float64 else ctypes.c_float(x)


def accel(pos, mass, soft2, dtype=np.float64):
    """acceleration_kernel_blocked_ (nbody.cu:19-65); pos is (N,3)."""
    pos = _c(pos, dtype); mass = _c(mass, dtype); soft2 = _c(soft2, dtype)
    out = np.empty_like(pos)
    getattr(lib(), f"orc_accel_{_suf(dtype)}")(_i64(len(mass)), _p(pos), _p(mass), _p(soft2), _p(out))
    return out


def accel_subset(pos, mass, soft2, idx, dtype=np.float64):
    """acceleration_kernel_blocked_ evaluated for the targets ``idx`` only."""
    pos = _c(pos, dtype); mass = _c(mass, dtype); soft2 = _c(soft2, dtype)
    idx = _c(idx, np.int64)
    out = np.empty((len(idx), 3), dtype=dtype)
    getattr(lib(), f"orc_accel_subset_{_suf(dtype)}")(
        _i64(len(mass)), _p(pos), _p(mass), _p(soft2), _i64(len(idx)), _p(idx), _p(out))
    return out


def grav(pos, mass, soft2, dtype=np.float64):
    """grav_kernel_blocked_ (nbody.cu:67-114), self term included."""
    pos = _c(pos, dtype); mass = _c(mass, dtype); soft2 = _c(soft2, dtype)
    out = np.empty(len(mass), dtype=dtype)
    getattr(lib(), f"orc_grav_{_suf(dtype)}")(_i64(len(mass)), _p(pos), _p(mass), _p(soft2), _p(out))
    return out


def predict(acc, vel, pos, dt, dtype=np.float64):
    """predict_step_ (nbody.cu:122-152). Returns (new_pos, vel_pred)."""
    acc = _c(acc, dtype); vel = _c(vel, dtype); pos = _c(pos, dtype).copy()
    vp = np.empty_like(vel)
    getattr(lib(), f"orc_predict_{_suf(dtype)}")(_i64(len(pos)), _p(acc), _p(vel), _p(pos), _p(vp), _r(dt, dtype))
    return pos, vp


def correct(acc, vel, vel_pred, dt, dtype=np.float64):
    """correct_step_ (nbody.cu:159-181). Returns the new velocities."""
    acc = _c(acc, dtype); vel = _c(vel, dtype).copy(); vp = _c(vel_pred, dtype)
    getattr(lib(), f"orc_correct_{_suf(dtype)}")(_i64(len(vel)), _p(acc), _p(vel), _p(vp), _r(dt, dtype))
    return vel


def conservation(pos, vel, mass, gr, dtype=np.float64):
    """calc_conservation (nbody.cu:246-291) -> dict(Ek, Ep, P, L)."""
    pos = _c(pos, dtype); vel = _c(vel, dtype); mass = _c(mass, dtype); gr = _c(gr, dtype)
    out = np.empty(8, dtype=dtype)
    getattr(lib(), f"orc_conservation_{_suf(dtype)}")(_i64(len(mass)), _p(pos), _p(vel), _p(mass), _p(gr), _p(out))
    return {"Ek": float(out[0]), "Ep": float(out[1]), "P": out[2:5].astype(np.float64),
            "L": out[5:8].astype(np.float64)}


def cic_float(pos, mass, nx, dx, bc_frac=0.0, dtype=np.float64):
    """deposit_cell_mass_CIC_ (rm_p2mesh.cu:176-236), float adds in particle order."""
    pos = _c(pos, dtype); mass = _c(mass, dtype)
    grid = np.empty(nx ** 3, dtype=dtype)
    getattr(lib(), f"orc_cic_float_{_suf(dtype)}")(_i64(len(mass)), _p(pos), _p(mass), _int(nx),
                                                    _r(dx, dtype), _r(bc_frac, dtype), _p(grid))
    return grid


def cic_fixed(pos, mass, nx, dx, e, bc_frac=0.0, dtype=np.float64):
    """CIC deposit with int64 fixed-point accumulation (resolution 2**-e)."""
    pos = _c(pos, dtype); mass = _c(mass, dtype)
    grid = np.empty(nx ** 3, dtype=np.int64)
    getattr(lib(), f"orc_cic_fixed_{_suf(dtype)}")(_i64(len(mass)), _p(pos), _p(mass), _int(nx),
                                                    _r(dx, dtype), _r(bc_frac, dtype), _int(e), _p(grid))
    return grid


def fixed_to_real(grid, e, dtype=np.float64):
    grid = _c(grid, np.int64)
    out = np.empty(len(grid), dtype=dtype)
    getattr(lib(), f"orc_fixed_to_real_{_suf(dtype)}")(_i64(len(grid)), _p(grid), _int(e), _p(out))
    return out


def wave_dt(dx, c, dtype=np.float64):
    """CFL step of validate_dependent_fields (io_json.cpp:174-178)."""
    f = getattr(lib(), f"orc_wave_dt_{_suf(dtype)}")
    return f(dx, c)


def rho_mean(total_mass, nx, dx, dtype=np.float64):
    """set_boundary_periodic's mean density (wave.cu:214-228)."""
    return getattr(lib(), f"orc_rho_mean_{_suf(dtype)}")(total_mass, nx, dx)


def wave_coef(diss, c, dt, dx, dtype=np.float64):
    out = np.empty(4, dtype=dtype)
    getattr(lib(), f"orc_wave_coef_{_suf(dtype)}")(_r(diss, dtype), _r(c, dtype), _r(dt, dtype), _r(dx, dtype), _p(out))
    return {"k_main": out[0], "k_f": out[1], "k_actual": out[2], "k_old": out[3]}


def wave_iteration(nx, dx, diss, c, dt, periodic, rho_mean_, phi, phi_old, mass, dtype=np.float64):
    """wave_diss_iteration (wave.cu:106-193). Returns phi_new."""
    phi = _c(phi, dtype); phi_old = _c(phi_old, dtype); mass = _c(mass, dtype)
    out = np.empty_like(phi)
    r = lambda x: _r(x, dtype)  # noqa: E731
    getattr(lib(), f"orc_wave_iteration_{_suf(dtype)}")(_int(nx), r(dx), r(diss), r(c), r(dt), _int(periodic),
                                                         r(rho_mean_), _p(out), _p(phi), _p(phi_old), _p(mass))
    return out


def wave_iteration_bc(nx, dx, diss, c, dt, boundary, phi, phi_old, mass, *, bc_frac=0.0, diss_extra=0.0,
                      rho_mean_=0.0, bc_mass=1.0, com=(0.0, 0.0, 0.0), r_min=None, dtype=np.float64):
    """wave_diss_iteration (wave.cu:106-193) for any boundary type, with the
    absorbing layer (bc_frac, diss_extra) and the monopole parameters."""
    phi = _c(phi, dtype); phi_old = _c(phi_old, dtype); mass = _c(mass, dtype)
    out = np.empty_like(phi)
    r = lambda x: _r(x, dtype)  # noqa: E731
    com = np.ascontiguousarray(com, dtype=dtype)
    rmin = 0.5 * dx if r_min is None else r_min
    getattr(lib(), f"orc_wave_iteration_bc_{_suf(dtype)}")(
        _int(nx), r(dx), r(bc_frac), r(diss), r(diss_extra), r(c), r(dt), _int(_bc(True, boundary)), r(rho_mean_),
        r(bc_mass), _p(com), r(rmin), _p(out), _p(phi), _p(phi_old), _p(mass))
    return out


def solve_wave(nx, dx, diss, c, dt, periodic, rho_mean_, prev, curr, mass, max_iter, tol,
               check_step=30, hold=3, dtype=np.float64):
    """solve_wave_equation (wave.cu:230-297). Returns (prev, curr, iterations)."""
    b = [_c(prev, dtype).copy(), _c(curr, dtype).copy(), np.zeros_like(_c(curr, dtype))]
    mass = _c(mass, dtype)
    perm = (ctypes.c_int * 3)()
    r = lambda x: _r(x, dtype)  # noqa: E731
    it = getattr(lib(), f"orc_solve_wave_{_suf(dtype)}")(
        _int(nx), r(dx), r(diss), r(c), r(dt), _int(periodic), r(rho_mean_),
        _p(b[0]), _p(b[1]), _p(b[2]), _p(mass), _int(max_iter), r(tol), _int(check_step), _int(hold), perm)
    return b[perm[0]], b[perm[1]], int(it)


def steady_err(u_new, u_old, dtype=np.float64):
    u_new = _c(u_new, dtype); u_old = _c(u_old, dtype)
    return getattr(lib(), f"orc_steady_err_{_suf(dtype)}")(_i64(len(u_new)), _p(u_new), _p(u_old))


def gradient(nx, dx, phi, dtype=np.float64):
    """calc_acceleration_field_ (rm_p2mesh.cu:238-290). Returns (nx^3, 3)."""
    phi = _c(phi, dtype)
    out = np.empty((nx ** 3, 3), dtype=dtype)
    getattr(lib(), f"orc_gradient_{_suf(dtype)}")(_int(nx), _r(dx, dtype), _p(phi), _p(out))
    return out


def ngp_bin(pos, nx, dx, bc_frac=0.0, dtype=np.float64):
    """NGP counting sort (rm_p2mesh.cu:25-126, with the intended exclusive
    prefix).  Returns (cell_of, count, start, order)."""
    pos = _c(pos, dtype)
    n = len(pos)
    cell_of = np.empty(n, dtype=np.int32); order = np.empty(n, dtype=np.int32)
    count = np.empty(nx ** 3, dtype=np.int32); start = np.empty(nx ** 3, dtype=np.int32)
    getattr(lib(), f"orc_ngp_bin_{_suf(dtype)}")(_i64(n), _p(pos), _int(nx), _r(dx, dtype), _r(bc_frac, dtype),
                                                _p(cell_of), _p(count), _p(start), _p(order))
    return cell_of, count, start, order


def ngp_project(nx, count, start, order, grav, acc, dtype=np.float64):
    """Cell averages of particle potential and acceleration (rm_p2mesh.cu:139-168)."""
    count = np.ascontiguousarray(count, np.int32); start = np.ascontiguousarray(start, np.int32)
    order = np.ascontiguousarray(order, np.int32)
    grav = _c(grav, dtype); acc = _c(acc, dtype)
    phi = np.empty(nx ** 3, dtype=dtype); cacc = np.empty((nx ** 3, 3), dtype=dtype)
    getattr(lib(), f"orc_ngp_project_{_suf(dtype)}")(_int(nx), _p(count), _p(start), _p(order), _p(grav), _p(acc),
                                                    _p(phi), _p(cacc))
    return phi, cacc


def gather(pos, nx, dx, field, bc_frac=0.0, dtype=np.float64):
    """interpolate_acceleration_to_particles_kernel (rm_mesh2p.cu:48-114)."""
    pos = _c(pos, dtype); field = _c(field, dtype)
    out = np.empty_like(pos)
    getattr(lib(), f"orc_gather_{_suf(dtype)}")(_i64(len(pos)), _p(pos), _int(nx), _r(dx, dtype),
                                                 _r(bc_frac, dtype), _p(field), _p(out))
    return out


def ngp_phi(pos, nx, dx, phi, bc_frac=0.0, dtype=np.float64):
    """NGP phi projection (rm_p2mesh.cu:34-53 + rm_mesh2p.cu:29-38)."""
    pos = _c(pos, dtype); phi = _c(phi, dtype)
    out = np.empty(len(pos), dtype=dtype)
    getattr(lib(), f"orc_ngp_phi_{_suf(dtype)}")(_i64(len(pos)), _p(pos), _int(nx), _r(dx, dtype),
                                                  _r(bc_frac, dtype), _p(phi), _p(out))
    return out


def nbody_run(pos, vel, mass, soft2, dt, nsteps, dtype=np.float64):
    """N-body strategy: precalc + nsteps x (predict, accel, correct).

    Returns (pos, vel, acc, grav_at_t0)."""
    pos = _c(pos, dtype).copy(); vel = _c(vel, dtype).copy()
    mass = _c(mass, dtype); soft2 = _c(soft2, dtype)
    acc = np.empty_like(pos); gr = np.empty(len(mass), dtype=dtype)
    getattr(lib(), f"orc_nbody_run_{_suf(dtype)}")(_i64(len(mass)), _p(pos), _p(vel), _p(mass), _p(soft2),
                                                    _r(dt, dtype), _int(nsteps), _p(acc), _p(gr))
    return pos, vel, acc, gr


def pm_run(pos, vel, mass, nsteps, *, nx, dx, dt, diss_base, wave_speed, setup_iterations, iterations,
           tolerance=0.0, check_step=30, hold=3, bc_frac=0.0, periodic=True, fixed_e=61, dtype=np.float64,
           boundary=None, diss_extra=0.0):
    """PM strategy: precalc + sync + nsteps x (predict, deposit, wave, gradient, gather, correct).
    ``boundary`` (name or BoundaryType value) overrides ``periodic``.

    Returns dict(pos, vel, acc, phi, phi_prev, grid_acc, grav, iterations)."""
    cf = PmConfig(nx=nx, dx=dx, bc_frac=bc_frac, diss_base=diss_base, wave_speed=wave_speed,
                  tolerance=tolerance, setup_iterations=setup_iterations, iterations=iterations,
                  check_step=check_step, hold=hold, boundary=_bc(periodic, boundary), fixed_e=fixed_e, dt=dt,
                  diss_extra=diss_extra)
    pos = _c(pos, dtype).copy(); vel = _c(vel, dtype).copy(); mass = _c(mass, dtype)
    n = len(mass)
    ncell = nx ** 3
    acc = np.empty_like(pos)
    phi2 = np.empty(2 * ncell, dtype=dtype)
    gacc = np.empty((ncell, 3), dtype=dtype)
    gr = np.empty(n, dtype=dtype)
    iters = np.zeros(nsteps + 1, dtype=np.int32)
    getattr(lib(), f"orc_pm_run_{_suf(dtype)}")(ctypes.byref(cf), _i64(n), _p(pos), _p(vel), _p(mass),
                                                 _int(nsteps), _p(acc), _p(phi2), _p(gacc), _p(gr), _p(iters))
    return {"pos": pos, "vel": vel, "acc": acc, "phi": phi2[:ncell], "phi_prev": phi2[ncell:],
            "grid_acc": gacc, "grav": gr, "iterations": iters}


def pm_adjoint(pos, vel, mass, nsteps, gx_T, gv_T=None, *, nx, dx, dt, diss_base, wave_speed, setup_iterations,
               iterations, tolerance=0.0, check_step=30, hold=3, bc_frac=0.0, periodic=True, fixed_e=61,
               boundary=None, diss_extra=0.0):
    """Reverse-mode gradient of pm_run (FP64): returns (dL/dx0, dL/dv0, dL/dm)
    for a loss with dL/dx_T = gx_T and dL/dv_T = gv_T.  Periodic and
    zero-Dirichlet boundaries only."""
    cf = PmConfig(nx=nx, dx=dx, bc_frac=bc_frac, diss_base=diss_base, wave_speed=wave_speed,
                  tolerance=tolerance, setup_iterations=setup_iterations, iterations=iterations,
                  check_step=check_step, hold=hold, boundary=_bc(periodic, boundary), fixed_e=fixed_e, dt=dt,
                  diss_extra=diss_extra)
    pos = _c(pos, np.float64); vel = _c(vel, np.float64); mass = _c(mass, np.float64)
    n = len(mass)
    gx = np.zeros((n, 3)) if gx_T is None else _c(gx_T, np.float64)
    gv = np.zeros((n, 3)) if gv_T is None else _c(gv_T, np.float64)
    gx0 = np.empty((n, 3)); gv0 = np.empty((n, 3)); gm = np.empty(n)
    rc = lib().orc_pm_adjoint_f64(ctypes.byref(cf), _i64(n), _p(pos), _p(vel), _p(mass), _int(nsteps), _p(gx),
                                  _p(gv), _p(gx0), _p(gv0), _p(gm))
    if rc != 0:
        raise ValueError("pm_adjoint: periodic and zero-Dirichlet boundaries only")
    return gx0, gv0, gm


def center_of_mass(pos, mass, fixed_e):
    """Order-independent fixed-point centre of mass used by the monopole boundaries.  FP64."""
    pos = _c(pos, np.float64); mass = _c(mass, np.float64)
    out = np.empty(3)
    lib().orc_center_of_mass_f64(_i64(len(mass)), _p(pos), _p(mass), _int(fixed_e), _p(out))
    return out


def gradient_t(gbar, nx, dx):
    """Transpose of gradient(): phibar = (-1/dx) D^T gbar, gbar of shape (nx^3, 3).  FP64."""
    gbar = _c(gbar, np.float64)
    out = np.empty(nx ** 3)
    lib().orc_gradient_t_f64(_int(nx), ctypes.c_double(dx), _p(gbar), _p(out))
    return out


def nbody_adjoint(pos, vel, mass, soft2, dt, nsteps, gx_T, gv_T=None):
    """Reverse-mode gradient of nbody_run: returns (dL/dx0, dL/dv0, dL/dm)
    for a loss with dL/dx_T = gx_T and dL/dv_T = gv_T.  FP64."""
    pos = _c(pos, np.float64); vel = _c(vel, np.float64)
    mass = _c(mass, np.float64); soft2 = _c(soft2, np.float64)
    n = len(mass)
    gx = _c(gx_T, np.float64)
    gv = np.zeros((n, 3)) if gv_T is None else _c(gv_T, np.float64)
    gx0 = np.empty((n, 3)); gv0 = np.empty((n, 3)); gm = np.empty(n)
    lib().orc_nbody_adjoint_f64(_i64(n), _p(pos), _p(vel), _p(mass), _p(soft2), ctypes.c_double(dt),
                                _int(nsteps), _p(gx), _p(gv), _p(gx0), _p(gv0), _p(gm))
    return gx0, gv0, gm


def vjp(pos, mass, soft2, u):
    """(J^T u, dm) of the direct-sum acceleration at pos.  FP64."""
    pos = _c(pos, np.float64); mass = _c(mass, np.float64); soft2 = _c(soft2, np.float64); u = _c(u, np.float64)
    g = np.empty_like(pos); gm = np.empty(len(mass))
    lib().orc_vjp_f64(_i64(len(mass)), _p(pos), _p(mass), _p(soft2), _p(u), _p(g), _p(gm))
    return g, gm
