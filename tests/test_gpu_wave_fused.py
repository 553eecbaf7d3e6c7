"""Several stencil iterations per launch (up to four at 128^3 and FP32 256^3,
three on the other periodic grids, two at most for FP64 512^3; shorter
launches for the remainders): bitwise the same as the
oracle's sequence of wave_diss_iteration calls (reference wave.cu:106-193,
solve_wave_equation wave.cu:230-297) for iteration counts with every
remainder, inside full PM runs (graph-replayed steps, the steady-state check
splitting the unchecked stretches, the recomputing adjoint), and bitwise
equal to the same runs with one iteration per launch.
"""
import numpy as np
import pytest

from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu

C = 4574.337022617616


def _fields(nx, dtype, seed):
    rng = np.random.default_rng(seed)
    ncell = nx ** 3
    prev = (rng.standard_normal(ncell) * 0.1).astype(dtype)
    curr = prev + (rng.standard_normal(ncell) * 0.01).astype(dtype)
    mass = (rng.uniform(0, 1, ncell) / ncell).astype(dtype)
    return prev, curr, mass


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nx,iters", [(64, 1), (64, 2), (64, 3), (64, 7), (64, 8), (128, 8), (128, 6), (128, 4),
                                     (128, 7), (128, 9), (256, 4), (256, 5), (256, 7)])
def test_wave_fused_iterations_bitwise(rrg, orc, dtype, nx, iters):
    dx = 16.0 / nx
    prev, curr, mass = _fields(nx, dtype, nx + iters)
    wdt = orc.wave_dt(dx, C, dtype=dtype)
    rm = orc.rho_mean(1.0, nx, dx, dtype=dtype)
    rp, rc, it = orc.solve_wave(nx, dx, 0.785, C, wdt, True, rm, prev, curr, mass, iters, 0.0, dtype=dtype)
    assert it == iters
    gp, gc = rrg.wave_iterate(nx, dx, 0.785, C, wdt, rm, iters, prev, curr, mass, dtype=dtype)
    assert np.array_equal(gp, rp) and np.array_equal(gc, rc)


@pytest.mark.parametrize("dtype,iters", [(np.float32, 5), (np.float64, 3)])
def test_wave_fused_512_bitwise(rrg, orc, dtype, iters):
    """FP32: one three- and one two-iteration launch; FP64: one two-iteration
    launch and a single one."""
    nx = 512
    dx = 16.0 / nx
    prev, curr, mass = _fields(nx, dtype, 5)
    wdt = orc.wave_dt(dx, C, dtype=dtype)
    rm = orc.rho_mean(1.0, nx, dx, dtype=dtype)
    rp, rc, _ = orc.solve_wave(nx, dx, 0.785, C, wdt, True, rm, prev, curr, mass, iters, 0.0, dtype=dtype)
    gp, gc = rrg.wave_iterate(nx, dx, 0.785, C, wdt, rm, iters, prev, curr, mass, dtype=dtype)
    assert np.array_equal(gp, rp) and np.array_equal(gc, rc)


def _run(rrg, dtype, *, nx, iters, setup, tol=0.0, check=30, hold=3, steps=3, fused=True, graphs=True, n=20_000):
    pos, vel, m = ics.exp_disk(n, seed=4)
    wave = dict(diss_base=0.785, diss_extra=0.01, wave_speed=C, setup_iterations=setup, iterations=iters,
                tolerance=tol, steady_check_step=check, steady_hold_steps=hold)
    real = "fp32" if dtype == np.float32 else "fp64"
    p = rrg.make_params(dt=5e-3, grid=dict(nx=nx, dx=16.0 / nx), wave=wave, real=real, wave_fused=fused,
                        graphs=graphs)
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        sim.sync_particles_and_grid()
        sim.calc_step(steps)
        sim.sync_particles_and_grid()
        return (pos, vel, m), sim.state(), sim.grid(), sim.stats()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("iters,setup", [(7, 9), (8, 12), (6, 10)])
def test_pm_run_fused_matches_oracle(rrg, orc, dtype, iters, setup):
    (pos, vel, m), st, gr, stats = _run(rrg, dtype, nx=64, iters=iters, setup=setup)
    o = orc.pm_run(pos, vel, m, 3, nx=64, dx=0.25, dt=5e-3, diss_base=0.785, wave_speed=C, setup_iterations=setup,
                   iterations=iters, fixed_e=stats["fixed_exponent"], dtype=dtype)
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(st[k].astype(dtype), o[k]), k
    assert np.array_equal(gr["phi"].astype(dtype), o["phi"])


def test_pm_steady_check_fused_matches_oracle(rrg, orc):
    """Checks every 5 iterations: fused launches fill the unchecked stretches
    (3 + 1 after the first check), the checking iteration is a single launch,
    the early exit is the oracle's."""
    (pos, vel, m), st, gr, stats = _run(rrg, np.float64, nx=64, iters=60, setup=400, tol=2e-3, check=5, hold=2,
                                        steps=2)
    o = orc.pm_run(pos, vel, m, 2, nx=64, dx=0.25, dt=5e-3, diss_base=0.785, wave_speed=C, setup_iterations=400,
                   iterations=60, tolerance=2e-3, check_step=5, hold=2, fixed_e=stats["fixed_exponent"])
    assert o["iterations"][0] < 400
    assert stats["last_wave_iterations"] == int(o["iterations"][-1])
    assert np.array_equal(st["pos"], o["pos"]) and np.array_equal(gr["phi"], o["phi"])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("check", [5, 9])
def test_pm_steady_check_fused_128_matches_oracle(rrg, orc, dtype, check):
    """128^3: the unchecked stretches between checks run four-iteration
    launches (4, and 4 + 4), then the single checking iteration; the early
    exit and the states are the oracle's, bitwise."""
    (pos, vel, m), st, gr, stats = _run(rrg, dtype, nx=128, iters=60, setup=400, tol=2e-3, check=check, hold=2,
                                        steps=2)
    o = orc.pm_run(pos, vel, m, 2, nx=128, dx=0.125, dt=5e-3, diss_base=0.785, wave_speed=C, setup_iterations=400,
                   iterations=60, tolerance=2e-3, check_step=check, hold=2, fixed_e=stats["fixed_exponent"],
                   dtype=dtype)
    assert o["iterations"][0] < 400
    assert stats["last_wave_iterations"] == int(o["iterations"][-1])
    assert np.array_equal(st["pos"].astype(dtype), o["pos"]) and np.array_equal(gr["phi"].astype(dtype), o["phi"])


@pytest.mark.parametrize("graphs", [True, False])
@pytest.mark.parametrize("dtype,nx,iters", [(np.float32, 128, 23), (np.float32, 256, 5), (np.float32, 256, 9),
                                            (np.float64, 128, 9), (np.float64, 512, 5)])
def test_fused_equals_single(rrg, graphs, dtype, nx, iters):
    """PM runs of the disk on the headline grid (23 = 7 x 3 + 2 iterations per
    step in FP32), 256^3 (5 = 3 + 2) and, in FP64, 128^3 and 512^3: fused
    launches on and off give the same bits.  The FP32 256^3 and FP64 shapes
    skip the source rows without mass (about 90% of the rows of the disk),
    which the single-iteration launches read."""
    out = []
    for fused in (True, False):
        _, st, gr, _ = _run(rrg, dtype, nx=nx, iters=iters, setup=31, steps=2, fused=fused, graphs=graphs)
        out.append((st, gr))
    (a, ga), (b, gb) = out
    assert np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["vel"], b["vel"])
    assert np.array_equal(ga["phi"], gb["phi"])


def test_fused_equals_single_every_row_with_mass(rrg):
    """FP64 128^3 with particles spread uniformly over the box, so every
    source row holds mass and the fused stencil reads every row: fused and
    single-iteration runs give the same bits."""
    rng = np.random.default_rng(12)
    n = 300_000
    pos = rng.uniform(-8.0, 8.0, (n, 3))
    vel = rng.normal(0.0, 0.1, (n, 3))
    m = np.full(n, 1.0 / n)
    out = []
    for fused in (True, False):
        wave = dict(diss_base=0.785, diss_extra=0.01, wave_speed=C, setup_iterations=13, iterations=9,
                    tolerance=0.0)
        p = rrg.make_params(dt=5e-3, grid=dict(nx=128, dx=0.125), wave=wave, real="fp64", wave_fused=fused)
        with rrg.Simulation(p, pos, vel, m) as sim:
            sim.precalc()
            sim.calc_step(2)
            sim.sync_particles_and_grid()
            out.append((sim.state(), sim.grid()))
    (a, ga), (b, gb) = out
    assert np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["vel"], b["vel"])
    assert np.array_equal(ga["phi"], gb["phi"])


def test_pm_adjoint_fused_equals_single(rrg):
    """The PM adjoint recomputes forward segments through the same schedule;
    the forward states and the gradients are bitwise equal (the adjoint's
    grid sums are exact integer sums)."""
    pos, vel, m = ics.exp_disk(2000, seed=2)
    rng = np.random.default_rng(5)
    gx, gv = rng.standard_normal((2000, 3)), rng.standard_normal((2000, 3))
    wave = dict(diss_base=0.785, diss_extra=0.01, wave_speed=C, setup_iterations=11, iterations=5, tolerance=0.0)
    out = []
    for fused in (True, False):
        p = rrg.make_params(dt=0.02, grid=dict(nx=64, dx=0.25), wave=wave, wave_fused=fused)
        with rrg.Simulation(p, pos, vel, m) as sim:
            sim.precalc()
            sim.adjoint_begin(3)
            sim.calc_step(7)
            st = sim.state()
            out.append((st["pos"], st["vel"]) + tuple(sim.adjoint(gx, gv)))
    a, b = out
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for ga, gb in zip(a[2:], b[2:]):
        assert np.array_equal(ga, gb)


@pytest.mark.parametrize("iters,setup", [(5, 11), (6, 9)])
def test_pm_adjoint_fused_matches_oracle(rrg, orc, iters, setup):
    """The transposed stencil runs three (and two) iterations per launch on
    64^3 with s-bar accumulated in the single-launch order; the gradients
    match the oracle's reverse sweep (the interpolation adjoint is summed in
    fixed point on the GPU, in FP64 on the CPU)."""
    n, steps = 1500, 7
    pos, vel, m = ics.exp_disk(n, seed=8)
    rng = np.random.default_rng(9)
    gx, gv = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    wave = dict(diss_base=0.785, diss_extra=0.0, wave_speed=C, setup_iterations=setup, iterations=iters, tolerance=0.0)
    p = rrg.make_params(dt=0.02, grid=dict(nx=64, dx=0.25), wave=wave)
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        sim.adjoint_begin(3)
        sim.calc_step(steps)
        gx0, gv0, gm = sim.adjoint(gx, gv)
        e = sim.stats()["fixed_exponent"]
    ox0, ov0, om = orc.pm_adjoint(pos, vel, m, steps, gx, gv, nx=64, dx=0.25, dt=0.02, diss_base=0.785, wave_speed=C,
                                  setup_iterations=setup, iterations=iters, fixed_e=e)
    rel = lambda a, b: np.max(np.abs(a - b)) / np.max(np.abs(b))  # noqa: E731
    assert rel(gx0, ox0) < 1e-9 and rel(gv0, ov0) < 1e-9 and rel(gm, om) < 1e-9


@pytest.mark.parametrize("iters,setup", [(5, 7), (6, 8)])
def test_pm_adjoint_fused_128_matches_oracle(rrg, orc, iters, setup):
    """128^3 FP64: the transposed stencil's three- and two-iteration shapes
    (consecutive cells per lane, input plane in LDS) against the oracle's
    reverse sweep."""
    n, steps = 3000, 4
    pos, vel, m = ics.exp_disk(n, seed=11)
    rng = np.random.default_rng(12)
    gx, gv = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    wave = dict(diss_base=0.785, diss_extra=0.0, wave_speed=C, setup_iterations=setup, iterations=iters, tolerance=0.0)
    p = rrg.make_params(dt=0.02, grid=dict(nx=128, dx=0.125), wave=wave)
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        sim.adjoint_begin(2)
        sim.calc_step(steps)
        gx0, gv0, gm = sim.adjoint(gx, gv)
        e = sim.stats()["fixed_exponent"]
    ox0, ov0, om = orc.pm_adjoint(pos, vel, m, steps, gx, gv, nx=128, dx=0.125, dt=0.02, diss_base=0.785,
                                  wave_speed=C, setup_iterations=setup, iterations=iters, fixed_e=e)
    rel = lambda a, b: np.max(np.abs(a - b)) / np.max(np.abs(b))  # noqa: E731
    assert rel(gx0, ox0) < 1e-9 and rel(gv0, ov0) < 1e-9 and rel(gm, om) < 1e-9


@pytest.mark.parametrize("nx,real", [(64, "fp32"), (128, "fp32"), (256, "fp32"), (512, "fp32"), (256, "fp64"),
                                     (512, "fp64")])
def test_pm_adjoint_fused_equals_single_large(rrg, nx, real):
    """The transposed stencil at the fused FP32 shapes and the FP64 shapes with
    several waves per row (three- and two-iteration launches): forward states
    and gradients bitwise equal to one iteration per launch (the adjoint's
    grid sums are exact integer sums)."""
    n = 4000
    pos, vel, m = ics.exp_disk(n, seed=3)
    rng = np.random.default_rng(4)
    gx, gv = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    wave = dict(diss_base=0.785, diss_extra=0.01, wave_speed=C, setup_iterations=7, iterations=5, tolerance=0.0)
    out = []
    for fused in (True, False):
        p = rrg.make_params(dt=0.02, grid=dict(nx=nx, dx=16.0 / nx), wave=wave, real=real, wave_fused=fused)
        with rrg.Simulation(p, pos, vel, m) as sim:
            sim.precalc()
            sim.adjoint_begin(2)
            sim.calc_step(4)
            st = sim.state()
            out.append((st["pos"], st["vel"]) + tuple(sim.adjoint(gx, gv)))
    a, b = out
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for ga, gb in zip(a[2:], b[2:]):
        assert np.array_equal(ga, gb), (nx, real, np.max(np.abs(ga - gb)) / np.max(np.abs(gb)))
