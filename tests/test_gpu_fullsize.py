"""The workloads at their stated sizes on the GPU, against the CPU oracle.

* PM headline: exponential disk N = 1,000,000 (seed 1), FP32 (and FP64), 128^3 cells on
  [-8, 8]^3, 200 unchecked damped-wave iterations per step, dt = 5e-3, 18
  steps with the particle arrays re-sorted by cell every 16 steps (the run
  crosses one re-sort; the default interval is 48, the same code).  Positions, velocities, accelerations and the
  potential must be bitwise those of oracle.pm_run (rm_p2mesh.cu:176-290,
  rm_mesh2p.cu:48-114, wave.cu:106-297), with the whole step replayed from one
  graph (as timed) and with one graph per stage.
* The 64-galaxy ensemble: N_g = round(2^U(12, 16)) (seed 2), Plummer / disk
  alternating, eps = 0.01, mixed FP32 forces on FP64 state, autotuned launch
  shape.  Forces at t = 0 on 64 sampled targets of every system against the
  FP64 direct sum (oracle.accel_subset, nbody.cu:19-65): median relative
  error <= 1e-6 per system; and 5 steps of the four smallest systems against
  independent oracle runs (nbody.cu:122-181).
* The CIC deposit's two paths on pre-sorted dense clusters: a workgroup whose
  256 particles' CIC stencils fit a box of at most 2048 cells accumulates in
  LDS and adds each box cell once; larger boxes issue each run's atomics
  directly.  Which path each workgroup takes is recomputed on the host from
  the kernel's rule, so the test states that both are reached, and that in
  FP32 on grids above 64 cells particles clamped onto the upper face (whose
  upper corner is their base cell) reach the box path; the grid must be
  bitwise oracle.cic_fixed, in sorted and in shuffled order.
"""
import numpy as np
import pytest

from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu

C = 4574.337022617616


def _pm_headline(rrg, timing, steps, real="fp32"):
    pos, vel, m = ics.exp_disk(1_000_000, seed=1)
    wave = dict(diss_base=0.785, diss_extra=0.01, wave_speed=C, setup_iterations=200, iterations=200,
                tolerance=0.0)
    p = rrg.make_params(dt=5e-3, grid=dict(nx=128, dx=0.125, bc_frac=0.0), wave=wave, real=real, sort_interval=16)
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        sim.sync_particles_and_grid()
        sim.set_timing(timing)
        sim.calc_step(steps)
        sim.synchronize()
        st = sim.state()
        phi = sim.grid()["phi"]
        stats = sim.stats()
    return pos, vel, m, st, phi, stats


@pytest.fixture(scope="module")
def headline_oracle(orc):
    pos, vel, m = ics.exp_disk(1_000_000, seed=1)
    return orc.pm_run(pos, vel, m, 18, nx=128, dx=0.125, dt=5e-3, diss_base=0.785, wave_speed=C,
                      setup_iterations=200, iterations=200, tolerance=0.0, fixed_e=60, dtype=np.float32)


@pytest.mark.parametrize("timing", [False, True])
def test_pm_headline_workload_bitwise(rrg, headline_oracle, timing):
    print(f"PM headline N=1e6 128^3 fp32, 18 steps, stage timing {timing}", flush=True)
    _, _, _, st, phi, stats = _pm_headline(rrg, timing, 18)
    o = headline_oracle
    assert stats["fixed_exponent"] == 60 and stats["steps"] == 18 and stats["last_wave_iterations"] == 200
    assert stats["graph_launches"] == (18 * 4 if timing else 18)  # deposit, source, wave, gather
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(st[k].astype(np.float32), o[k]), k
    assert np.array_equal(phi.astype(np.float32), o["phi"])


def test_pm_headline_workload_fp64_bitwise(rrg, orc):
    """The headline workload in FP64 (the reference's default precision):
    the four-iteration FP64 stencil reads only the source rows that hold
    mass (the disk leaves ~90% of them empty) and must still give the
    oracle's bits after 18 steps across a re-sort."""
    print("PM headline N=1e6 128^3 fp64, 18 steps", flush=True)
    pos, vel, m, st, phi, stats = _pm_headline(rrg, False, 18, real="fp64")
    o = orc.pm_run(pos, vel, m, 18, nx=128, dx=0.125, dt=5e-3, diss_base=0.785, wave_speed=C,
                   setup_iterations=200, iterations=200, tolerance=0.0, fixed_e=stats["fixed_exponent"],
                   dtype=np.float64)
    assert stats["steps"] == 18 and stats["graph_launches"] == 18
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(st[k], o[k]), k
    assert np.array_equal(phi, o["phi"])


def test_ensemble_64_galaxies(rrg, orc):
    print("ensemble: 64 galaxies seed 2, mixed, autotuned", flush=True)
    sizes = ics.ensemble_sizes(64, seed=2)
    rng = np.random.default_rng(2)
    parts = [ics.plummer(int(nn), rng=rng) if g % 2 == 0 else ics.exp_disk(int(nn), rng=rng)
             for g, nn in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    pos = np.concatenate([q[0] for q in parts]); vel = np.concatenate([q[1] for q in parts])
    m = np.concatenate([q[2] for q in parts])
    s2 = np.full(len(m), 1e-4)
    p = rrg.make_params(dt=1e-3, force="mixed", autotune=True)
    with rrg.Simulation(p, pos, vel, m, s2, offsets=off) as sim:
        sim.precalc()
        sim.sync_particles_and_grid()
        acc0 = sim.state()["acc"]
        sim.calc_step(5)
        st = sim.state()
        stats = sim.stats()
    assert stats["systems"] == 64 and stats["tuned"] == 1 and stats["graph_launches"] >= 1
    srng = np.random.default_rng(7)
    worst = 0.0
    for g in range(64):
        a, b = off[g], off[g + 1]
        idx = np.sort(srng.choice(b - a, 64, replace=False))
        ref = orc.accel_subset(pos[a:b], m[a:b], s2[a:b], idx)
        err = np.linalg.norm(acc0[a:b][idx] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        worst = max(worst, float(np.median(err)))
        assert np.median(err) <= 1e-6 and np.max(err) < 1e-4, (g, np.median(err), np.max(err))
    print(f"worst per-system median force error {worst:.3e}", flush=True)
    for g in np.argsort(sizes)[:4]:
        a, b = off[g], off[g + 1]
        opos, ovel, _, _ = orc.nbody_run(pos[a:b], vel[a:b], m[a:b], s2[a:b], 1e-3, 5)
        assert np.max(np.abs(st["pos"][a:b] - opos)) < 1e-7, g
        assert np.max(np.abs(st["vel"][a:b] - ovel)) < 1e-6, g


def _cic_cells(pos, nx, dx, dtype):
    """Base cell of the CIC stencil as the deposit computes it (clamped
    offsets; FP32 arithmetic for FP32 state), its key and the upper corner."""
    dt = dtype
    dmin = dt(-0.5 * (dt(dx) * nx))
    inv_dx = dt(np.float32(1.0) / np.float32(dt(dx)))
    upper = dt(dt(nx - 1) - dt(1e-6))
    u = np.clip((pos.astype(dt) - dmin) * inv_dx - dt(0.5), dt(0), upper)
    b = np.clip(np.floor(u).astype(np.int64), 0, nx - 1)
    key = b[:, 0] + nx * (b[:, 1] + nx * b[:, 2])
    top = np.minimum(b + 1, nx - 1)
    return key, b, top


def _path_counts(b, top, box=2048):
    """Workgroups of 256 particles on the LDS-box path (bounding box of their
    stencils within `box` cells) and on the direct path, and the clamped
    particles (upper corner = base cell) on the box path."""
    nbox = ndirect = clamped_in_box = 0
    for w0 in range(0, len(b), 256):
        vol = int(np.prod(top[w0:w0 + 256].max(axis=0) - b[w0:w0 + 256].min(axis=0) + 1))
        if vol <= box:
            nbox += 1
            clamped_in_box += int(np.sum(np.any(top[w0:w0 + 256] == b[w0:w0 + 256], axis=1)))
        else:
            ndirect += 1
    return nbox, ndirect, clamped_in_box


def _clusters(nx, dx, dtype, seed=4):
    rng = np.random.default_rng(seed)
    half = nx * dx / 2
    parts = [
        rng.normal(0.0, 1.5 * dx, (30_000, 3)),                              # dense centre: long runs, table
        rng.uniform(-half, half, (20_000, 3)),                               # spread: keys far apart, direct
        half + rng.uniform(0.0, 2 * dx, (6_000, 3)),                         # beyond the upper corner
        np.column_stack([half + rng.uniform(0, dx, 6_000), rng.normal(0, dx, (6_000, 2))]),  # upper x face
        -half - rng.uniform(0.0, 2 * dx, (6_000, 3)),                        # beyond the lower corner
    ]
    pos = np.concatenate(parts).astype(dtype)
    m = (rng.uniform(0.5, 1.5, len(pos)) / len(pos)).astype(dtype)
    return pos, m


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cic_merge_paths_bitwise(rrg, orc, dtype):
    nx, dx, e = 128, 0.125, 60
    pos, m = _clusters(nx, dx, dtype)
    key, b, top = _cic_cells(pos, nx, dx, dtype)
    order = np.argsort(key, kind="stable")
    pos, m, b, top = pos[order], m[order], b[order], top[order]
    nbox, ndirect, clamped = _path_counts(b, top)
    print(f"cic {np.dtype(dtype).name}: workgroups on the LDS box path {nbox}, direct {ndirect}, clamped particles "
          f"in boxes {clamped}", flush=True)
    assert nbox > 100 and ndirect > 20
    if dtype == np.float32:
        assert clamped > 0  # clamped onto the last cell: upper corner = base cell
    ref = orc.cic_fixed(pos, m, nx, dx, e, dtype=dtype)
    assert np.array_equal(rrg.cic_deposit(pos, m, nx, dx, e, dtype=dtype), ref)
    perm = np.random.default_rng(1).permutation(len(m))
    assert np.array_equal(rrg.cic_deposit(pos[perm], m[perm], nx, dx, e, dtype=dtype), ref)


def _gather_boxes(pos, nx, dx, dtype, box):
    """Workgroups of 256 particles (in cell order) whose gather corners fit a
    box of at most `box` cells (the gather's LDS path) and the others (each
    particle forms its corners' -grad phi itself); the corners as the gather
    computes them ((x - min) / dx - 0.5, floored, each index clamped)."""
    dt = dtype
    dmin = dt(-0.5 * (dt(dx) * nx))
    f = np.floor(((pos.astype(dt) - dmin) / dt(dx) - dt(0.5)).astype(np.float32)).astype(np.int64)
    lo, hi = np.clip(f, 0, nx - 1), np.clip(f + 1, 0, nx - 1)
    nbox = nfall = 0
    for w0 in range(0, len(pos), 256):
        vol = int(np.prod(hi[w0:w0 + 256].max(axis=0) - lo[w0:w0 + 256].min(axis=0) + 1))
        if vol <= box:
            nbox += 1
        else:
            nfall += 1
    return nbox, nfall


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gather_phi_paths_bitwise(rrg, orc, dtype):
    """The steps' gather forms -grad phi itself: workgroups whose corner cells
    fit the LDS box fill it once, the others form each particle's corners
    directly.  Clusters that reach both paths (and particles beyond the box
    faces): the accelerations after precalc + sync and the state after two
    kicked steps are the oracle's (k_gradient followed by k_gather), bitwise."""
    nx, dx = 128, 0.125
    pos, m = _clusters(nx, dx, dtype, seed=7)
    vel = np.zeros_like(pos)
    key, _, _ = _cic_cells(pos, nx, dx, dtype)
    order = np.argsort(key, kind="stable")  # the library's particle order after precalc's sort
    box = 1024 if dtype == np.float32 else 512
    nbox, nfall = _gather_boxes(pos[order], nx, dx, dtype, box)
    print(f"gather {np.dtype(dtype).name}: workgroups on the LDS box path {nbox}, per-particle {nfall}", flush=True)
    assert nbox > 100 and nfall > 20
    real = "fp32" if dtype == np.float32 else "fp64"
    wave = dict(diss_base=0.785, diss_extra=0.01, wave_speed=C, setup_iterations=20, iterations=8, tolerance=0.0)
    p = rrg.make_params(dt=5e-3, grid=dict(nx=nx, dx=dx, bc_frac=0.0), wave=wave, real=real)
    with rrg.Simulation(p, pos.astype(np.float64), vel, m.astype(np.float64)) as sim:
        sim.precalc()
        sim.sync_particles_and_grid()
        acc0 = sim.state()["acc"]
        e = sim.stats()["fixed_exponent"]
        sim.calc_step(2)
        st = sim.state()
    kw = dict(nx=nx, dx=dx, dt=5e-3, diss_base=0.785, wave_speed=C, setup_iterations=20, iterations=8,
              tolerance=0.0, fixed_e=e, dtype=dtype)
    o0 = orc.pm_run(pos, vel, m, 0, **kw)
    assert np.array_equal(acc0.astype(dtype), o0["acc"])
    o2 = orc.pm_run(pos, vel, m, 2, **kw)
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(st[k].astype(dtype), o2[k]), k
