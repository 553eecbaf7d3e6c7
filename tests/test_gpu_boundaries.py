"""Non-periodic boundaries on the GPU (zero Dirichlet, monopole Dirichlet,
monopole radiative, all with the absorbing layer; reference wave.cu:18-193)
against the oracle, bitwise: positions, velocities, accelerations, the
potential and its gradient after precalc + steps, including the steady-check
path; and the PM adjoint with the absorbing layer against the oracle gradient.
"""
import math

import numpy as np
import pytest

from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu

C = 4574.337022617616
BOUNDARIES = ["zero_dirichlet", "monopole_dirichlet", "monopole_radiative"]


def _run(rrg, orc, boundary, dtype, *, tol=0.0, steps=4, setup=40, iters=12, bc_frac=0.25, diss_extra=1.5):
    pos, vel, m = ics.exp_disk(8000, seed=3)
    wave = dict(diss_base=0.785, diss_extra=diss_extra, wave_speed=C, setup_iterations=setup, iterations=iters,
                tolerance=tol, steady_check_step=5, steady_hold_steps=2)
    real = "fp32" if dtype == np.float32 else "fp64"
    p = rrg.make_params(dt=5e-3, grid=dict(nx=16, dx=1.0, bc_frac=bc_frac), wave=wave, real=real, boundary=boundary)
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        sim.sync_particles_and_grid()
        sim.calc_step(steps)
        sim.sync_particles_and_grid()
        st = sim.state()
        gr = sim.grid()
        stats = sim.stats()
    o = orc.pm_run(pos, vel, m, steps, nx=16, dx=1.0, dt=5e-3, diss_base=0.785, wave_speed=C,
                   setup_iterations=setup, iterations=iters, tolerance=tol, check_step=5, hold=2, bc_frac=bc_frac,
                   boundary=boundary, diss_extra=diss_extra, fixed_e=stats["fixed_exponent"], dtype=dtype)
    return st, gr, stats, o


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("boundary", BOUNDARIES)
def test_boundary_pm_run_bitwise(rrg, orc, boundary, dtype):
    st, gr, _, o = _run(rrg, orc, boundary, dtype)
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(st[k].astype(dtype), o[k]), (boundary, k)
    assert np.array_equal(gr["phi"].astype(dtype), o["phi"]), boundary
    assert np.array_equal(gr["acc"].astype(dtype), o["grid_acc"]), boundary
    if boundary == "monopole_dirichlet":  # the faces hold -1/r about the centre of mass
        assert np.min(gr["phi"]) < 0


@pytest.mark.parametrize("boundary", BOUNDARIES)
def test_boundary_steady_check_bitwise(rrg, orc, boundary):
    st, gr, stats, o = _run(rrg, orc, boundary, np.float64, tol=2e-3, setup=600, iters=80, steps=2)
    assert stats["last_wave_iterations"] == int(o["iterations"][-1])
    assert o["iterations"][0] < 600
    assert np.array_equal(st["pos"], o["pos"]) and np.array_equal(gr["phi"], o["phi"])


def test_pm_adjoint_with_absorbing_layer_matches_oracle(rrg, orc):
    n, steps = 500, 7
    pos, vel, m = ics.exp_disk(n, seed=2)
    rng = np.random.default_rng(5)
    gx, gv = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    wave = dict(diss_base=0.785, diss_extra=4.0, wave_speed=C, setup_iterations=20, iterations=6, tolerance=0.0)
    p = rrg.make_params(dt=0.05, grid=dict(nx=16, dx=1.0, bc_frac=0.5), wave=wave, boundary="zero_dirichlet")
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        sim.adjoint_begin(3)
        sim.calc_step(steps)
        gx0, gv0, gm = sim.adjoint(gx, gv)
        e = sim.stats()["fixed_exponent"]
    ox0, ov0, om = orc.pm_adjoint(pos, vel, m, steps, gx, gv, nx=16, dx=1.0, dt=0.05, diss_base=0.785, wave_speed=C,
                                  setup_iterations=20, iterations=6, bc_frac=0.5, diss_extra=4.0,
                                  boundary="zero_dirichlet", fixed_e=e)
    rel = lambda a, b: np.max(np.abs(a - b)) / np.max(np.abs(b))  # noqa: E731
    assert rel(gx0, ox0) < 1e-9 and rel(gv0, ov0) < 1e-9 and rel(gm, om) < 1e-9


@pytest.mark.parametrize("boundary", ["monopole_dirichlet", "monopole_radiative"])
def test_pm_adjoint_rejects_monopole(rrg, boundary):
    pos, vel, m = ics.exp_disk(300, seed=1)
    wave = dict(diss_base=0.785, diss_extra=0.0, wave_speed=C, setup_iterations=5, iterations=3, tolerance=0.0)
    p = rrg.make_params(dt=0.05, grid=dict(nx=16, dx=1.0), wave=wave, boundary=boundary)
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        with pytest.raises(rrg.RRGError, match="periodic and zero-Dirichlet"):
            sim.adjoint_begin(2)


def test_monopole_centre_of_mass_is_order_free(rrg, orc):
    """The monopole centre is the fixed-point centre of mass: the GPU run of a
    shifted disk matches the oracle, and permuting the particles changes
    nothing (bitwise)."""
    pos, vel, m = ics.exp_disk(3000, seed=9)
    pos = pos + np.array([0.7, -0.4, 0.2])
    wave = dict(diss_base=0.785, diss_extra=1.0, wave_speed=C, setup_iterations=30, iterations=10, tolerance=0.0)
    p = rrg.make_params(dt=5e-3, grid=dict(nx=16, dx=1.0, bc_frac=0.2), wave=wave, boundary="monopole_dirichlet")
    out = []
    perm = np.random.default_rng(1).permutation(len(m))
    for idx in (np.arange(len(m)), perm):
        with rrg.Simulation(p, pos[idx], vel[idx], m[idx]) as sim:
            sim.precalc()
            out.append(sim.grid()["phi"])
    assert np.array_equal(out[0], out[1])
    e = 61 - math.frexp(float(np.sum(m)))[1]
    com = orc.center_of_mass(pos, m, e)
    assert np.max(np.abs(com - np.average(pos, axis=0, weights=m))) < 1e-12


def test_monopole_precalc_again_after_set_state(rrg):
    """A second precalc recomputes the centre of mass; steps replayed from
    captured graphs must use it like eagerly launched steps do (the grids are
    not reset by precalc, so both runs carry the same history)."""
    pos, vel, m = ics.exp_disk(4000, seed=12)
    shifted = pos + np.array([0.9, -0.6, 0.3])
    wave = dict(diss_base=0.785, diss_extra=1.0, wave_speed=C, setup_iterations=30, iterations=10, tolerance=0.0)
    out = []
    for graphs in (True, False):
        p = rrg.make_params(dt=5e-3, grid=dict(nx=16, dx=1.0, bc_frac=0.2), wave=wave, boundary="monopole_dirichlet",
                            graphs=graphs)
        with rrg.Simulation(p, pos, vel, m) as sim:
            sim.precalc()
            sim.calc_step(3)
            sim.set_state(pos=shifted, vel=vel)
            sim.precalc()
            sim.sync_particles_and_grid()
            sim.calc_step(3)
            out.append(sim.state())
    assert np.array_equal(out[0]["pos"], out[1]["pos"]) and np.array_equal(out[0]["vel"], out[1]["vel"])
