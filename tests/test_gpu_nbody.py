"""Direct-sum path on the GPU vs the CPU oracle (through the C ABI).

Oracle: oracle/rrg_oracle.cpp restating nbody.cu:19-181 (acceleration,
potential, predictor/corrector) and nbody.cu:246-291 (conservation).
Tolerances: FP64 kernel vs FP64 oracle -- summation order and fused
multiply-adds differ, so agreement is to a few ulps of the per-particle
force scale; mixed kernel -- median relative force error <= 1e-6 (the
target bound), p99 reported.
"""
import numpy as np
import pytest

from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu


def _rel(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)


@pytest.mark.parametrize("force", ["fp64", "fp64_exact"])
@pytest.mark.parametrize("n", [1, 2, 255, 1000, 4096])
def test_fp64_accel_matches_oracle(rrg, orc, n, force):
    pos, _, m = ics.plummer(max(n, 2), seed=0)
    pos, m = pos[:n], m[:n]
    s2 = np.full(n, 1e-4)
    ref = orc.accel(pos, m, s2)
    got = rrg.direct_accel(pos, m, s2, force=force)
    scale = np.sqrt(np.mean(np.sum(ref ** 2, axis=1))) + 1e-300
    err = np.max(np.linalg.norm(got - ref, axis=1)) / scale
    print(f"n={n} fp64 max err / rms|a| = {err:.3e}")
    assert err < 1e-13


@pytest.mark.parametrize("force,tol", [("fp64", 1e-14), ("fp64_exact", 1e-14), ("mixed", 1e-6), ("fp32", 1e-6)])
def test_accel_unequal_softening(rrg, orc, force, tol):
    """Non-uniform softening takes the per-pair 0.5*(s_i + s_j) path."""
    pos, _, m = ics.plummer(777, seed=5)
    rng = np.random.default_rng(5)
    s2 = rng.uniform(1e-5, 1e-3, len(m))
    m = m * rng.uniform(0.5, 1.5, len(m))
    ref = orc.accel(pos, m, s2)
    got = rrg.direct_accel(pos, m, s2, force=force)
    assert np.median(_rel(got, ref)) < tol
    if force != "fp64":
        return
    grav_ref = orc.grav(pos, m, s2)
    grav = rrg.direct_grav(pos, m, s2, force="fp64")
    assert np.max(np.abs(grav - grav_ref) / np.abs(grav_ref)) < 1e-13


@pytest.mark.parametrize("n", [3, 1000, 4096])
def test_fp64_grav_matches_oracle(rrg, orc, n):
    pos, _, m = ics.plummer(n, seed=1)
    s2 = np.full(n, 1e-4)
    ref = orc.grav(pos, m, s2)
    got = rrg.direct_grav(pos, m, s2, force="fp64")
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-13


def test_mixed_accel_error_c3_sample(rrg, orc):
    """Disk N=196,608 + bulge N=65,536, eps=0.005: median relative force error
    of the mixed kernel on 4096 sampled particles vs the FP64 oracle."""
    pos, _, m = ics.disk_bulge()
    n = len(m)
    s2 = np.full(n, 0.005 ** 2)
    idx = np.sort(np.random.default_rng(7).choice(n, 4096, replace=False))
    ref = orc.accel_subset(pos, m, s2, idx)
    got = rrg.direct_accel(pos, m, s2, force="mixed")[idx]
    rel = _rel(got, ref)
    med, p99, mx = np.median(rel), np.quantile(rel, 0.99), np.max(rel)
    print(f"mixed C3 sample: median {med:.3e} p99 {p99:.3e} max {mx:.3e}")
    assert med <= 1e-6 and p99 <= 1e-6
    rel32 = _rel(rrg.direct_accel(pos, m, s2, force="fp32")[idx], ref)
    print(f"fp32 C3 sample: median {np.median(rel32):.3e} p99 {np.quantile(rel32, 0.99):.3e} max {rel32.max():.3e}")
    assert np.median(rel32) <= 1e-6
    for force in ("fp64", "fp64_exact"):
        got64 = rrg.direct_accel(pos, m, s2, force=force)[idx]
        assert np.median(_rel(got64, ref)) < 1e-13


@pytest.mark.parametrize("n", [1, 100, 4096])
def test_mixed_grav(rrg, orc, n):
    pos, _, m = ics.plummer(max(n, 2), seed=2)
    pos, m = pos[:n], m[:n]
    s2 = np.full(n, 1e-4)
    ref = orc.grav(pos, m, s2)
    got = rrg.direct_grav(pos, m, s2, force="mixed")
    assert np.median(np.abs(got - ref) / np.abs(ref)) < 1e-6


def _nbody_sim(rrg, pos, vel, m, s2, dt, force="fp64"):
    p = rrg.make_params(dt=dt, force=force)
    sim = rrg.Simulation(p, pos, vel, m, s2)
    sim.precalc()
    return sim


def test_c1_run_matches_oracle(rrg, orc):
    """Plummer N=4096, eps=0.01, dt=1e-3, 100 steps, FP64 (config C1)."""
    pos, vel, m = ics.plummer(4096, seed=0)
    s2 = np.full(len(m), 1e-4)
    sim = _nbody_sim(rrg, pos, vel, m, s2, 1e-3)
    c0 = sim.conservation()
    sim.calc_step(100)
    st = sim.state()
    c1 = sim.conservation()
    sim.close()
    opos, ovel, oacc, ograv0 = orc.nbody_run(pos, vel, m, s2, 1e-3, 100)
    dx = np.max(np.abs(st["pos"] - opos))
    dv = np.max(np.abs(st["vel"] - ovel))
    print(f"C1 100 steps: max|dx| {dx:.3e} max|dv| {dv:.3e}")
    assert dx < 1e-11 and dv < 1e-10
    assert np.max(np.abs(st["acc"] - oacc)) < 1e-10
    # energy: reference convention, self term included in the potential
    oc0 = orc.conservation(pos, vel, m, ograv0)
    ograv1 = orc.grav(opos, m, s2)
    oc1 = orc.conservation(opos, ovel, m, ograv1)
    for a, b in ((c0, oc0), (c1, oc1)):
        assert abs(a["Ek"] - b["Ek"]) < 1e-12
        assert abs(a["Ep"] - b["Ep"]) < 1e-12
        assert np.max(np.abs(a["L"] - b["L"])) < 1e-12
    drift = (c1["E"] / c0["E"]) - 1.0
    odrift = (oc1["Ek"] + oc1["Ep"]) / (oc0["Ek"] + oc0["Ep"]) - 1.0
    print(f"C1 energy drift gpu {drift:.6e} oracle {odrift:.6e}")
    assert abs(drift - odrift) < 1e-11


def test_mixed_run_tracks_fp64(rrg, orc):
    pos, vel, m = ics.plummer(4096, seed=0)
    s2 = np.full(len(m), 1e-4)
    sim = _nbody_sim(rrg, pos, vel, m, s2, 1e-3, force="mixed")
    sim.calc_step(20)
    st = sim.state()
    sim.close()
    opos, ovel, _, _ = orc.nbody_run(pos, vel, m, s2, 1e-3, 20)
    assert np.max(np.abs(st["pos"] - opos)) < 1e-8
    assert np.max(np.abs(st["vel"] - ovel)) < 1e-6


def test_step_before_precalc_is_an_error(rrg):
    pos, vel, m = ics.plummer(64, seed=0)
    p = rrg.make_params(dt=1e-3)
    with rrg.Simulation(p, pos, vel, m, np.full(64, 1e-4)) as sim:
        with pytest.raises(rrg.RRGError) as ei:
            sim.calc_step(1)
        assert ei.value.code == rrg._lib.RRG_ERR_STATE


def test_conservation_matches_numpy(rrg):
    pos, vel, m = ics.plummer(3000, seed=9)
    s2 = np.full(len(m), 1e-4)
    with _nbody_sim(rrg, pos, vel, m, s2, 1e-3) as sim:
        sim.sync_particles_and_grid()
        c = sim.conservation()
        g = sim.state()["grav"]
    ek = 0.5 * np.sum(m * np.sum(vel ** 2, axis=1))
    ep = 0.5 * np.sum(m * g)
    P = np.sum(m[:, None] * vel, axis=0)
    Lm = np.sum(m[:, None] * np.cross(pos, vel), axis=0)
    assert abs(c["Ek"] - ek) < 1e-13 and abs(c["Ep"] - ep) < 1e-13
    assert np.allclose(c["P"], P, atol=1e-15) and np.allclose(c["L"], Lm, atol=1e-15)


@pytest.mark.parametrize("force", ["fp64", "mixed", "fp32"])
@pytest.mark.parametrize("ensemble", [False, True])
def test_graph_steps_equal_eager_steps(rrg, force, ensemble):
    """19 steps from captured graphs (two graphs of eight steps, in which each
    step's corrector and the next step's predictor run as one launch, then
    three one-step graphs) against 19 eagerly launched steps: positions,
    velocities and accelerations bitwise, for one system and for an
    ensemble of three."""
    rng = np.random.default_rng(4)
    sizes = [1500, 700, 2048] if ensemble else [3000]
    parts = [ics.plummer(n, rng=rng) for n in sizes]
    pos = np.concatenate([p for p, _, _ in parts])
    vel = np.concatenate([v for _, v, _ in parts])
    m = np.concatenate([q for _, _, q in parts])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64) if ensemble else None
    s2 = np.full(len(m), 1e-4)
    out = []
    for graphs in (True, False):
        p = rrg.make_params(dt=1e-3, force=force, graphs=graphs)
        with rrg.Simulation(p, pos, vel, m, s2, offsets=off) as sim:
            sim.precalc()
            sim.calc_step(19)
            out.append(sim.state())
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(out[0][k], out[1][k]), k
