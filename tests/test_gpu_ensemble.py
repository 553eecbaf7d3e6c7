"""Ensembles (one persistent work-queue launch over uneven systems), hipGraph
replay and the autotuner, on the GPU, against the CPU oracle and against the
library's own single-system path.

Graph checks replay the captured step several times with eager work and
synchronisations in between and require every replay to match eager
execution bitwise (same kernels, same order).
"""
import numpy as np
import pytest

from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu

SIZES = [300, 1000, 2500, 4096, 777, 5000]


def _ensemble(sizes, seed=11):
    rng = np.random.default_rng(seed)
    parts = []
    for g, n in enumerate(sizes):
        if g % 2 == 0:
            p, v, m = ics.plummer(n, rng=rng)
        else:
            p, v, m = ics.exp_disk(n, rng=rng)
        parts.append((p, v, m))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    pos = np.concatenate([p for p, _, _ in parts]); vel = np.concatenate([v for _, v, _ in parts])
    m = np.concatenate([q for _, _, q in parts])
    return pos, vel, m, off


@pytest.mark.parametrize("force,tol", [("fp64", 1e-11), ("mixed", 1e-7), ("fp32", 1e-5)])
def test_ensemble_matches_independent_oracle_runs(rrg, orc, force, tol):
    pos, vel, m, off = _ensemble(SIZES)
    s2 = np.full(len(m), 1e-4)
    p = rrg.make_params(dt=1e-3, force=force)
    with rrg.Simulation(p, pos, vel, m, s2, offsets=off) as sim:
        sim.precalc()
        sim.calc_step(10)
        st = sim.state()
        cons = sim.conservation_systems()
        stats = sim.stats()
    assert stats["systems"] == len(SIZES) and stats["work_items"] > len(SIZES)
    for g in range(len(SIZES)):
        a, b = off[g], off[g + 1]
        opos, ovel, _, _ = orc.nbody_run(pos[a:b], vel[a:b], m[a:b], s2[a:b], 1e-3, 10)
        dx = np.max(np.abs(st["pos"][a:b] - opos))
        dv = np.max(np.abs(st["vel"][a:b] - ovel))
        assert dx < tol and dv < 10 * tol, (g, dx, dv)
        # per-system conservation = numpy sums over that system
        ek = 0.5 * np.sum(m[a:b] * np.sum(st["vel"][a:b] ** 2, axis=1))
        assert abs(cons[g]["Ek"] - ek) < 1e-12
        L = np.sum(m[a:b, None] * np.cross(st["pos"][a:b], st["vel"][a:b]), axis=0)
        assert np.allclose(cons[g]["L"], L, atol=1e-14)


def test_ensemble_potential_matches_oracle(rrg, orc):
    pos, vel, m, off = _ensemble(SIZES[:3])
    s2 = np.full(len(m), 1e-4)
    p = rrg.make_params(dt=1e-3)
    with rrg.Simulation(p, pos, vel, m, s2, offsets=off) as sim:
        sim.precalc()
        sim.sync_particles_and_grid()
        g = sim.state()["grav"]
    for k in range(3):
        a, b = off[k], off[k + 1]
        ref = orc.grav(pos[a:b], m[a:b], s2[a:b])
        assert np.max(np.abs(g[a:b] - ref) / np.abs(ref)) < 1e-13


@pytest.mark.parametrize("force", ["fp64", "mixed"])
def test_autotuned_shape_gives_same_physics(rrg, force):
    pos, vel, m, off = _ensemble(SIZES)
    s2 = np.full(len(m), 1e-4)
    out = []
    for tune in (False, True):
        p = rrg.make_params(dt=1e-3, force=force, autotune=tune)
        with rrg.Simulation(p, pos, vel, m, s2, offsets=off) as sim:
            sim.precalc()
            sim.calc_step(3)
            out.append(sim.state())
            stats = sim.stats()
    assert stats["tuned"] == 1
    tol = 1e-12 if force == "fp64" else 1e-6
    assert np.max(np.abs(out[0]["pos"] - out[1]["pos"])) < tol


@pytest.mark.parametrize("ensemble", [False, True])
@pytest.mark.parametrize("force", ["fp64", "mixed"])
@pytest.mark.parametrize("timing", [False, True])
def test_graph_replay_bitwise_equals_eager(rrg, ensemble, force, timing):
    """Steps replayed from graphs equal eager steps: one graph per step
    without stage timing, one per stage (3) with it."""
    pos, vel, m, off = _ensemble(SIZES if ensemble else [4096])
    s2 = np.full(len(m), 1e-4)
    states = {}
    for graphs in (False, True):
        p = rrg.make_params(dt=1e-3, force=force, graphs=graphs)
        with rrg.Simulation(p, pos, vel, m, s2, offsets=off if ensemble else None) as sim:
            sim.precalc()
            sim.set_timing(timing)
            seq = []
            for _ in range(4):
                sim.calc_step(2)
                sim.conservation()  # eager work and a synchronisation between replays
                seq.append(sim.state())
            states[graphs] = (seq, sim.stats()["graph_launches"])
    assert states[False][1] == 0 and states[True][1] == (24 if timing else 8)
    for a, b in zip(states[False][0], states[True][0]):
        for k in ("pos", "vel", "acc"):
            assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("ensemble", [False, True])
def test_multi_step_graphs_bitwise_equal_eager(rrg, ensemble):
    """Runs of 19 steps without timing or recording: two graphs of 8 steps
    and three one-step graphs, bitwise the eager steps."""
    pos, vel, m, off = _ensemble(SIZES if ensemble else [2048])
    s2 = np.full(len(m), 1e-4)
    out = {}
    for graphs in (False, True):
        p = rrg.make_params(dt=1e-3, force="mixed", graphs=graphs)
        with rrg.Simulation(p, pos, vel, m, s2, offsets=off if ensemble else None) as sim:
            sim.precalc()
            sim.calc_step(19)
            st = sim.state()
            out[graphs] = (st, sim.stats()["graph_launches"], sim.stats()["steps"], sim.stats()["time"])
    assert out[True][1] == 5 and out[False][1] == 0
    assert out[True][2] == out[False][2] == 19 and out[True][3] == out[False][3]
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(out[False][0][k], out[True][0][k]), k


@pytest.mark.parametrize("nx", [32, 64])
@pytest.mark.parametrize("timing", [False, True])
def test_pm_graph_replay_bitwise_equals_eager(rrg, nx, timing):
    """PM steps from graphs equal eager steps (64^3: fused stencil launches
    and four potential grids); one graph per step without stage timing, one
    per stage (4: deposit, source, wave, gather) with it."""
    pos, vel, m = ics.exp_disk(20_000, seed=1)
    wave = dict(diss_base=0.785, diss_extra=0.01, wave_speed=4574.337022617616, setup_iterations=30,
                iterations=7, tolerance=0.0)
    res = {}
    for graphs in (False, True):
        p = rrg.make_params(dt=5e-3, grid=dict(nx=nx, dx=16.0 / nx), wave=wave, real="fp32", graphs=graphs,
                            sort_interval=8)
        with rrg.Simulation(p, pos, vel, m) as sim:
            sim.precalc()
            sim.set_timing(timing)
            seq = []
            for _ in range(6):  # crosses a re-sort and all three buffer rotations
                sim.calc_step(3)
                sim.conservation()
                seq.append((sim.state(), sim.grid()["phi"]))
            res[graphs] = (seq, sim.stats()["graph_launches"])
    assert res[True][1] == (72 if timing else 18)
    for (a, pa), (b, pb) in zip(res[False][0], res[True][0]):
        assert np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["vel"], b["vel"])
        assert np.array_equal(pa, pb)


def test_ensemble_ragged_tiny_systems(rrg, orc):
    """Systems of 1, 2 and 3 particles next to larger ones (items shorter
    than one tile, single-particle systems with only the self pair)."""
    sizes = [1, 2, 3, 257, 64, 1, 700]
    rng = np.random.default_rng(21)
    parts = [ics.plummer(max(k, 2), rng=rng) for k in sizes]
    pos = np.concatenate([p[:k] for (p, _, _), k in zip(parts, sizes)])
    vel = np.concatenate([v[:k] for (_, v, _), k in zip(parts, sizes)])
    m = np.concatenate([q[:k] for (_, _, q), k in zip(parts, sizes)])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    s2 = np.full(len(m), 1e-3)
    for force, tol in (("fp64", 1e-12), ("mixed", 1e-7)):
        p = rrg.make_params(dt=1e-3, force=force)
        with rrg.Simulation(p, pos, vel, m, s2, offsets=off) as sim:
            sim.precalc()
            sim.calc_step(5)
            st = sim.state()
        for g in range(len(sizes)):
            a, b = off[g], off[g + 1]
            opos, ovel, _, _ = orc.nbody_run(pos[a:b], vel[a:b], m[a:b], s2[a:b], 1e-3, 5)
            assert np.max(np.abs(st["pos"][a:b] - opos)) < tol, (force, g)
            assert np.max(np.abs(st["vel"][a:b] - ovel)) < 10 * tol, (force, g)
