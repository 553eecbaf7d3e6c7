"""The steady-state early exit (solve_wave_equation with SteadyStateChecker,
wave.cu:253-287 and wave_steady_checker.cu:95-143) decided on the device, so
PM steps with tolerance > 0 stay graph-replayed: every step's iteration count
and the final fields against the oracle's run (oracle.pm_run restates the
checker and is pinned by tests/test_oracle_pins.py), with graphs (the
device-side decision: one graph per step, or one per stage with stage
timing) and without (the host reads each check back), on the fused
four- and three-iteration stencils, the single-iteration kernel and the
non-periodic kernels.  The cases mix steps that stop early with steps that
run to the iteration limit."""
import numpy as np
import pytest

from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu

C = 4574.337022617616
STEPS = 6

CASES = [
    # nx, dx, tolerance, real, boundary
    (128, 0.125, 3e-5, "fp32", "periodic"),
    (128, 0.125, 3e-5, "fp64", "periodic"),
    (64, 0.25, 1e-5, "fp64", "periodic"),
    (32, 0.5, 1e-5, "fp32", "periodic"),
    (32, 0.5, 1e-5, "fp64", "zero_dirichlet"),
    (24, 0.75, 1e-4, "fp64", "monopole_dirichlet"),
]


def _run(rrg, nx, dx, tol, real, boundary, graphs, timing=False, check=5, setup=400, iterations=60, steps=STEPS):
    pos, vel, m = ics.exp_disk(20_000, seed=1)
    wave = dict(diss_base=0.785, diss_extra=0.01, wave_speed=C, setup_iterations=setup, iterations=iterations,
                tolerance=tol, steady_check_step=check, steady_hold_steps=2)
    p = rrg.make_params(dt=5e-3, grid=dict(nx=nx, dx=dx, bc_frac=0.1 if boundary != "periodic" else 0.0),
                        wave=wave, real=real, boundary=boundary, graphs=graphs)
    iters = []
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        iters.append(sim.stats()["last_wave_iterations"])
        sim.sync_particles_and_grid()
        sim.set_timing(timing)
        for _ in range(steps):
            sim.calc_step(1)
            iters.append(sim.stats()["last_wave_iterations"])
        sim.sync_particles_and_grid()
        st = sim.state()
        phi = sim.grid()["phi"]
        stats = sim.stats()
    return (pos, vel, m), iters, st, phi, stats


@pytest.mark.parametrize("nx,dx,tol,real,boundary", CASES)
def test_device_checked_steps_match_oracle(rrg, orc, nx, dx, tol, real, boundary):
    print(f"steady check on the device: nx={nx} tol={tol} {real} {boundary}", flush=True)
    dtype = np.float32 if real == "fp32" else np.float64
    (pos, vel, m), it_g, st_g, phi_g, stats = _run(rrg, nx, dx, tol, real, boundary, graphs=True)
    assert stats["graph_launches"] == STEPS
    _, it_h, st_h, phi_h, stats_h = _run(rrg, nx, dx, tol, real, boundary, graphs=False)
    assert stats_h["graph_launches"] == 0
    o = orc.pm_run(pos, vel, m, STEPS, nx=nx, dx=dx, bc_frac=0.1 if boundary != "periodic" else 0.0, dt=5e-3,
                   diss_base=0.785, diss_extra=0.01, wave_speed=C, setup_iterations=400, iterations=60,
                   tolerance=tol, check_step=5, hold=2, fixed_e=stats["fixed_exponent"], dtype=dtype,
                   boundary=boundary)
    assert it_g == it_h == [int(k) for k in o["iterations"]], (it_g, it_h, list(o["iterations"]))
    for st, phi in ((st_g, phi_g), (st_h, phi_h)):
        for k in ("pos", "vel", "acc", "grav"):
            assert np.array_equal(st[k].astype(dtype), o[k]), k
        assert np.array_equal(phi.astype(dtype), o["phi"])


def test_device_check_mixes_early_and_full_solves(rrg, orc):
    """The headline grid case above has both kinds of steps (the oracle's
    counts, so the test states what it covers)."""
    pos, vel, m = ics.exp_disk(20_000, seed=1)
    o = orc.pm_run(pos, vel, m, STEPS, nx=128, dx=0.125, dt=5e-3, diss_base=0.785, wave_speed=C,
                   setup_iterations=400, iterations=60, tolerance=3e-5, check_step=5, hold=2, fixed_e=60,
                   dtype=np.float32)
    steps = [int(k) for k in o["iterations"][1:]]
    assert 60 in steps and any(k < 60 for k in steps), steps


@pytest.mark.parametrize("graphs", [True, False])
def test_device_check_pair_stretches(rrg, orc, graphs):
    """Checks every 13 iterations on the 128^3 FP32 grid: each stretch between
    checks is three four-iteration launches that pass (phi^n, phi^(n-1)) as
    pair grids, and a converged solve's remaining stretches (pairs
    included) return at once.  The oracle's counts mix full solves (60) with
    solves that stop at 53 and 40 iterations; counts, particles and phi
    bitwise."""
    tol = 5e-5
    (pos, vel, m), its, st, phi, stats = _run(rrg, 128, 0.125, tol, "fp32", "periodic", graphs=graphs, check=13)
    o = orc.pm_run(pos, vel, m, STEPS, nx=128, dx=0.125, dt=5e-3, diss_base=0.785, diss_extra=0.01, wave_speed=C,
                   setup_iterations=400, iterations=60, tolerance=tol, check_step=13, hold=2,
                   fixed_e=stats["fixed_exponent"], dtype=np.float32)
    steps = [int(k) for k in o["iterations"][1:]]
    assert 60 in steps and 40 in steps, steps
    assert its == [int(k) for k in o["iterations"]], (its, list(o["iterations"]))
    for k in ("pos", "vel", "acc", "grav"):
        assert np.array_equal(st[k].astype(np.float32), o[k]), k
    assert np.array_equal(phi.astype(np.float32), o["phi"])


@pytest.mark.parametrize("nx,check", [(256, 10), (512, 13)])
def test_device_check_large_fp64(rrg, orc, nx, check):
    """The FP64 shapes with several waves per row and source-row flags, whose
    checking launches are fused: 256^3 (three iterations per launch) and
    512^3 (two, as in the reference's sample workload), short solves (40
    setup iterations, two steps of at most 40) at a tolerance the oracle's
    solves meet after 21 (256^3) or 27 (512^3) iterations and the setup does
    not, so the fused checks decide both ways; counts and fields bitwise."""
    dx, tol = 16.0 / nx, 0.03
    print(f"steady check on the device: nx={nx} fp64 check every {check}", flush=True)
    (pos, vel, m), its, st, phi, stats = _run(rrg, nx, dx, tol, "fp64", "periodic", graphs=True, check=check,
                                              setup=40, iterations=40, steps=2)
    o = orc.pm_run(pos, vel, m, 2, nx=nx, dx=dx, dt=5e-3, diss_base=0.785, diss_extra=0.01, wave_speed=C,
                   setup_iterations=40, iterations=40, tolerance=tol, check_step=check, hold=2,
                   fixed_e=stats["fixed_exponent"], dtype=np.float64)
    want = [int(k) for k in o["iterations"]]
    assert want[0] == 40 and min(want[1:]) < 40, want
    assert its == want, (its, want)
    for k in ("pos", "vel", "acc", "grav"):
        assert np.array_equal(st[k], o[k]), k
    assert np.array_equal(phi, o["phi"])


def test_device_check_with_stage_graphs(rrg):
    """Stage timing on: the wave stage is its own graph; same counts and
    fields as the whole-step graphs."""
    a = _run(rrg, 128, 0.125, 3e-5, "fp32", "periodic", graphs=True, timing=True)
    b = _run(rrg, 128, 0.125, 3e-5, "fp32", "periodic", graphs=True, timing=False)
    assert a[4]["graph_launches"] == 4 * STEPS and a[1] == b[1]  # deposit, source, wave, gather
    for k in ("pos", "vel"):
        assert np.array_equal(a[2][k], b[2][k]), k
    assert np.array_equal(a[3], b[3])
