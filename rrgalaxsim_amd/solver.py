"""Python view of the solver context and operators of librrg.so.

``Simulation`` mirrors the reference's ``IntegrationSolver`` strategy
(solver/app/include/rrgsim_integration.cuh:22-42): ``precalc``, ``calc_step``,
``sync_particles_and_grid``; the model (direct sum or PM "wave") follows
``use_wave_model`` exactly as ``integrate()`` chooses it
(solver/app/src/rrgsim_integration.cu:41-47).  The operator functions mirror
the reference's L2 free functions (nbody.cuh, rm_p2mesh.cuh, rm_mesh2p.cuh,
wave.cuh).  All compute runs in the HIP library.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


def default_params() -> L.Params:
    p = L.Params()
    L.check(L.lib().rrg_params_default(ctypes.byref(p)))
    return p


def wave_schedule(nx, iterations, real="fp32", boundary="periodic", fused=True) -> dict:
    """{iterations per launch: launches} for `iterations` unchecked stencil
    iterations of a context with this grid (the library's own rule)."""
    out = (ctypes.c_int32 * 4)()
    L.check(L.lib().rrg_wave_schedule({"fp64": L.REAL_F64, "fp32": L.REAL_F32}[real], int(nx), L.BOUNDARIES[boundary],
                                      int(iterations), 1 if fused else 0, out))
    return {1: out[0], 2: out[1], 3: out[2], 4: out[3]}


def make_params(*, dt, time_max=None, dt_save=None, grid=None, wave=None, real="fp64", force="fp64",
                boundary="periodic", fixed_exponent=0, device=0, graphs=True, autotune=False,
                wave_fused=True, sort_interval=0) -> L.Params:
    """Build an ``rrg_params``.  ``grid`` = dict(nx, dx, bc_frac); ``wave`` =
    dict(diss_base, diss_extra, wave_speed[, setup_iterations, iterations,
    tolerance, steady_check_step, steady_hold_steps])."""
    p = default_params()
    p.sim.dt_dynamics = dt
    p.sim.time_max = time_max if time_max is not None else dt
    p.sim.dt_save = dt_save if dt_save is not None else p.sim.time_max
    if grid is not None:
        p.has_grid = 1
        p.grid.nx = int(grid["nx"])
        p.grid.dx = float(grid["dx"])
        p.grid.bc_frac = float(grid.get("bc_frac", 0.0))
    if wave is not None:
        p.has_wave = 1
        for k, v in wave.items():
            setattr(p.wave, k, v)
    p.real_type = {"fp64": L.REAL_F64, "fp32": L.REAL_F32}[real]
    p.force_kind = L.FORCE_KINDS[force]
    p.boundary = L.BOUNDARIES[boundary]
    p.fixed_exponent = fixed_exponent
    p.device = device
    p.use_graphs = 1 if graphs else 0
    p.autotune = 1 if autotune else 0
    p.wave_fused = 1 if wave_fused else 0
    p.sort_interval = int(sort_interval)
    L.check(L.lib().rrg_params_finalize(ctypes.byref(p)))
    return p


def parse_params_json(path: str):
    p = L.Params()
    ini = L.IniParams()
    L.check(L.lib().rrg_params_parse_json(path.encode(), ctypes.byref(p), ctypes.byref(ini)))
    return p, ini


def read_ini(ini: L.IniParams, base_dir: str | None = None):
    n = L._i64(0)
    bd = base_dir.encode() if base_dir else None
    L.check(L.lib().rrg_ini_read(ctypes.byref(ini), bd, ctypes.byref(n), None, None, None, None))
    N = n.value
    pos = np.empty((N, 3)); vel = np.empty((N, 3)); mass = np.empty(N); soft2 = np.empty(N)
    L.check(L.lib().rrg_ini_read(ctypes.byref(ini), bd, ctypes.byref(n), L.ptr(pos), L.ptr(vel), L.ptr(mass),
                                 L.ptr(soft2)))
    return pos, vel, mass, soft2


class Simulation:
    """One solver context on one GPU.

    ``offsets`` (k+1 entries) makes it an ensemble of k independent systems
    (direct-sum model): system g is particles offsets[g]:offsets[g+1]."""

    def __init__(self, params: L.Params, pos, vel, mass, soft2=None, offsets=None):
        self.n = len(mass)
        self.params = params
        pos = L.f64(pos, (self.n, 3)); vel = L.f64(vel, (self.n, 3)); mass = L.f64(mass)
        s2 = None if soft2 is None else L.f64(soft2)
        h = L._vp()
        if offsets is None:
            self.k = 1
            L.check(L.lib().rrg_create(ctypes.byref(params), self.n, L.ptr(pos), L.ptr(vel), L.ptr(mass), L.ptr(s2),
                                       ctypes.byref(h)))
        else:
            off = np.ascontiguousarray(offsets, dtype=np.int64)
            self.k = len(off) - 1
            L.check(L.lib().rrg_create_ensemble(ctypes.byref(params), self.k, L.ptr(off), L.ptr(pos), L.ptr(vel),
                                                L.ptr(mass), L.ptr(s2), ctypes.byref(h)))
        self._h = h

    # IntegrationSolver interface -----------------------------------------
    def precalc(self) -> None:
        L.check(L.lib().rrg_precalc(self._h))

    def calc_step(self, nsteps: int = 1) -> None:
        L.check(L.lib().rrg_step(self._h, int(nsteps)))

    def sync_particles_and_grid(self) -> None:
        L.check(L.lib().rrg_sync(self._h))

    # state ----------------------------------------------------------------
    def state(self):
        pos = np.empty((self.n, 3)); vel = np.empty((self.n, 3)); acc = np.empty((self.n, 3))
        grav = np.empty(self.n)
        L.check(L.lib().rrg_get_state(self._h, L.ptr(pos), L.ptr(vel), L.ptr(acc), L.ptr(grav)))
        return {"pos": pos, "vel": vel, "acc": acc, "grav": grav}

    def set_state(self, pos=None, vel=None) -> None:
        p = None if pos is None else L.f64(pos, (self.n, 3))
        v = None if vel is None else L.f64(vel, (self.n, 3))
        L.check(L.lib().rrg_set_state(self._h, L.ptr(p), L.ptr(v)))

    def get_state_dev(self, pos=0, vel=0, acc=0, stream=0) -> None:
        """Device-array form of state(): FP64 (n,3) rows at the given device
        addresses (0 = skip), ordered after the work queued on `stream`."""
        L.check(L.lib().rrg_get_state_dev(self._h, pos or None, vel or None, acc or None, stream or None))

    def set_state_dev(self, pos=0, vel=0, stream=0) -> None:
        """Device-array form of set_state()."""
        L.check(L.lib().rrg_set_state_dev(self._h, pos or None, vel or None, stream or None))

    def adjoint_dev(self, gx_T=0, gv_T=0, gx_0=0, gv_0=0, gm=0, stream=0) -> None:
        """Device-array form of adjoint(): seeds and gradients at device
        addresses (0 = none / skip)."""
        L.check(L.lib().rrg_adjoint_dev(self._h, gx_T or None, gv_T or None, gx_0 or None, gv_0 or None,
                                        gm or None, stream or None))

    def conservation(self) -> dict:
        c = L.Conservation()
        L.check(L.lib().rrg_get_conservation(self._h, ctypes.byref(c)))
        return {"time": c.time, "Ek": c.Ek, "Ep": c.Ep, "E": c.Ek + c.Ep,
                "P": np.array(c.momentum[:]), "L": np.array(c.angular[:])}

    def conservation_systems(self) -> list:
        """Per-system conservation sums (one dict per system)."""
        arr = (L.Conservation * self.k)()
        L.check(L.lib().rrg_get_conservation_systems(self._h, self.k, arr))
        return [{"time": c.time, "Ek": c.Ek, "Ep": c.Ep, "E": c.Ek + c.Ep, "P": np.array(c.momentum[:]),
                 "L": np.array(c.angular[:])} for c in arr]

    def grid(self) -> dict:
        nx = self.params.grid.nx
        m = np.empty(nx ** 3); phi = np.empty(nx ** 3); acc = np.empty((nx ** 3, 3))
        L.check(L.lib().rrg_get_grid(self._h, L.ptr(m), L.ptr(phi), L.ptr(acc)))
        return {"mass": m, "phi": phi, "acc": acc}

    def stats(self) -> dict:
        s = L.Stats()
        L.check(L.lib().rrg_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in L.Stats._fields_}

    def synchronize(self) -> None:
        L.check(L.lib().rrg_synchronize(self._h))

    def adjoint_begin(self, checkpoint_every: int = 20) -> None:
        """Record the following steps for rrg_adjoint (checkpoints every K steps)."""
        L.check(L.lib().rrg_adjoint_begin(self._h, int(checkpoint_every)))

    def adjoint(self, gx_T=None, gv_T=None):
        """Gradients (dL/dx0, dL/dv0, dL/dm) of the recorded run for the given
        dL/dx_T and dL/dv_T."""
        gx = None if gx_T is None else L.f64(gx_T, (self.n, 3))
        gv = None if gv_T is None else L.f64(gv_T, (self.n, 3))
        gx0 = np.empty((self.n, 3)); gv0 = np.empty((self.n, 3)); gm = np.empty(self.n)
        L.check(L.lib().rrg_adjoint(self._h, L.ptr(gx), L.ptr(gv), L.ptr(gx0), L.ptr(gv0), L.ptr(gm)))
        return gx0, gv0, gm

    def set_timing(self, enabled: bool) -> None:
        """Per-stage HIP-event timing on the context stream (resets totals)."""
        L.check(L.lib().rrg_set_timing(self._h, 1 if enabled else 0))

    def timing(self) -> dict:
        """{stage: (ms, launches)} accumulated since set_timing(True)."""
        t = L.Timing()
        L.check(L.lib().rrg_get_timing(self._h, ctypes.byref(t)))
        return {name: (t.ms[k], t.launches[k]) for k, name in enumerate(L.STAGES)}

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.check(L.lib().rrg_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- operators
def _rt(dtype) -> int:
    return L.REAL_F32 if np.dtype(dtype) == np.float32 else L.REAL_F64


def direct_accel(pos, mass, soft2, force="fp64"):
    """nbody_acceleration (nbody.cu:200-213) -> (N,3) float64."""
    n = len(mass)
    pos = L.f64(pos, (n, 3)); mass = L.f64(mass); soft2 = L.f64(soft2)
    out = np.empty((n, 3))
    kind = L.FORCE_KINDS[force]
    L.check(L.lib().rrg_op_direct_accel(kind, n, L.ptr(pos), L.ptr(mass), L.ptr(soft2), L.ptr(out)))
    return out


def direct_grav(pos, mass, soft2, force="fp64"):
    """nbody_grav (nbody.cu:215-228), self term included -> (N,) float64."""
    n = len(mass)
    pos = L.f64(pos, (n, 3)); mass = L.f64(mass); soft2 = L.f64(soft2)
    out = np.empty(n)
    kind = L.FORCE_KINDS[force]
    L.check(L.lib().rrg_op_direct_grav(kind, n, L.ptr(pos), L.ptr(mass), L.ptr(soft2), L.ptr(out)))
    return out


def cic_deposit(pos, mass, nx, dx, e, bc_frac=0.0, dtype=np.float64):
    """deposit_cell_mass_CIC_ with int64 fixed point -> (nx^3,) int64."""
    n = len(mass)
    pos = np.ascontiguousarray(pos, dtype=dtype).reshape(n, 3)
    mass = np.ascontiguousarray(mass, dtype=dtype)
    out = np.empty(nx ** 3, dtype=np.int64)
    L.check(L.lib().rrg_op_cic_deposit(_rt(dtype), n, L.ptr(pos), L.ptr(mass), nx, dx, bc_frac, e, L.ptr(out)))
    return out


def wave_iterate(nx, dx, diss_base, wave_speed, wave_dt, rho_mean, iterations, phi_prev, phi_curr, mass,
                 boundary="periodic", dtype=np.float64):
    """iterations x wave_diss_iteration with buffer rotation -> (phi_prev, phi_curr)."""
    prev = np.ascontiguousarray(phi_prev, dtype=dtype).copy()
    curr = np.ascontiguousarray(phi_curr, dtype=dtype).copy()
    mass = np.ascontiguousarray(mass, dtype=dtype)
    bc = {"periodic": L.BC_PERIODIC, "zero_dirichlet": L.BC_ZERO_DIRICHLET}[boundary]
    L.check(L.lib().rrg_op_wave_iterate(_rt(dtype), nx, dx, diss_base, wave_speed, wave_dt, bc, rho_mean,
                                        iterations, L.ptr(prev), L.ptr(curr), L.ptr(mass)))
    return prev, curr


def gradient(nx, dx, phi, dtype=np.float64):
    """calc_acceleration_field_ -> (nx^3, 3)."""
    phi = np.ascontiguousarray(phi, dtype=dtype)
    out = np.empty((nx ** 3, 3), dtype=dtype)
    L.check(L.lib().rrg_op_gradient(_rt(dtype), nx, dx, L.ptr(phi), L.ptr(out)))
    return out


def gather(pos, nx, dx, field, bc_frac=0.0, dtype=np.float64):
    """interpolate_acceleration_to_particles_kernel -> (N, 3)."""
    n = len(pos)
    pos = np.ascontiguousarray(pos, dtype=dtype).reshape(n, 3)
    field = np.ascontiguousarray(field, dtype=dtype).reshape(nx ** 3, 3)
    out = np.empty((n, 3), dtype=dtype)
    L.check(L.lib().rrg_op_gather(_rt(dtype), n, L.ptr(pos), nx, dx, bc_frac, L.ptr(field), L.ptr(out)))
    return out


def ngp_bin(pos, nx, dx, bc_frac=0.0, dtype=np.float64):
    """NGP counting sort of particles by cell -> (cell_of, count, start, order) (int32)."""
    n = len(pos)
    pos = np.ascontiguousarray(pos, dtype=dtype).reshape(n, 3)
    cell_of = np.empty(n, np.int32); order = np.empty(n, np.int32)
    count = np.empty(nx ** 3, np.int32); start = np.empty(nx ** 3, np.int32)
    L.check(L.lib().rrg_op_ngp_bin(_rt(dtype), n, L.ptr(pos), nx, dx, bc_frac, L.ptr(cell_of), L.ptr(count),
                                   L.ptr(start), L.ptr(order)))
    return cell_of, count, start, order


def steady_err(u_new, u_old, dtype=np.float64) -> float:
    a = np.ascontiguousarray(u_new, dtype=dtype); b = np.ascontiguousarray(u_old, dtype=dtype)
    e = L._f64(0)
    L.check(L.lib().rrg_op_steady_err(_rt(dtype), len(a), L.ptr(a), L.ptr(b), ctypes.byref(e)))
    return e.value
