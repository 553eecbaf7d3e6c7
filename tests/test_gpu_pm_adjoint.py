"""Adjoint of the PM run on the GPU vs the FP64 CPU reverse-mode gradient of
the oracle (oracle pm_adjoint, itself pinned by finite differences of the
oracle's PM run in tests/test_oracle_pm_adjoint.py).

Bars: FP64 state agrees with the oracle gradient to 1e-9 relative (the
interpolation adjoint is summed in int64 fixed point, resolution
2^-E with n max|abar + h vbar| 2^E < 2^61, where the oracle adds in FP64); the
FP32 state's gradient is within 1e-4 of the FP64 reference, as
max |g - g_ref| / max |g_ref| per output.  Gradients repeat bitwise from run
to run (the fixed-point sums do not depend on the order of the atomics).
"""
import math

import numpy as np
import pytest

from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu

C = 4574.337022617616


def _wave(**kw):
    w = dict(diss_base=0.785, diss_extra=0.01, wave_speed=C, setup_iterations=20, iterations=6, tolerance=0.0)
    w.update(kw)
    return w


def _oracle_cfg(p_wave, nx, dx, dt, periodic):
    return dict(nx=nx, dx=dx, dt=dt, diss_base=p_wave["diss_base"], wave_speed=C,
                setup_iterations=p_wave["setup_iterations"], iterations=p_wave["iterations"],
                tolerance=p_wave["tolerance"], check_step=p_wave.get("steady_check_step", 30),
                hold=p_wave.get("steady_hold_steps", 3), periodic=periodic)


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _run(rrg, pos, vel, m, steps, every, gx, gv, *, real="fp64", boundary="periodic", nx=16, dx=1.0, dt=0.05,
         wave=None, graphs=True):
    p = rrg.make_params(dt=dt, grid=dict(nx=nx, dx=dx), wave=wave or _wave(), real=real, boundary=boundary,
                        graphs=graphs)
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        sim.adjoint_begin(every)
        sim.calc_step(steps)
        st = sim.state()
        g = sim.adjoint(gx, gv)
        after = sim.state()
        st["fixed_e"] = sim.stats()["fixed_exponent"]
    return st, g, after


@pytest.mark.parametrize("boundary", ["periodic", "zero_dirichlet"])
def test_pm_adjoint_fp64_matches_oracle(rrg, orc, boundary):
    n, steps = 500, 9
    pos, vel, m = ics.exp_disk(n, seed=2)
    rng = np.random.default_rng(5)
    gx, gv = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    st, (gx0, gv0, gm), _ = _run(rrg, pos, vel, m, steps, 4, gx, gv, boundary=boundary)  # uneven last segment
    cfg = _oracle_cfg(_wave(), 16, 1.0, 0.05, boundary == "periodic")
    o = orc.pm_run(pos, vel, m, steps, fixed_e=st["fixed_e"], **cfg)
    assert np.array_equal(st["pos"], o["pos"])  # the recorded forward is the plain forward
    ox0, ov0, om = orc.pm_adjoint(pos, vel, m, steps, gx, gv, fixed_e=st["fixed_e"], **cfg)
    print(f"{boundary}: rel err x0 {_rel(gx0, ox0):.2e} v0 {_rel(gv0, ov0):.2e} m {_rel(gm, om):.2e}")
    assert _rel(gx0, ox0) < 1e-9 and _rel(gv0, ov0) < 1e-9 and _rel(gm, om) < 1e-9


def test_pm_adjoint_fp32_within_1e4_of_fp64_reference(rrg, orc):
    n, steps = 2000, 10
    pos, vel, m = ics.exp_disk(n, seed=3)
    target = pos + 0.05
    wave = _wave(setup_iterations=200, iterations=30)
    p = rrg.make_params(dt=0.02, grid=dict(nx=32, dx=0.5), wave=wave, real="fp32")
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        sim.adjoint_begin(5)
        sim.calc_step(steps)
        xT = sim.state()["pos"]
        gxT = 2 * (xT - target) / xT.size
        gx0, gv0, gm = sim.adjoint(gxT)
        e = sim.stats()["fixed_exponent"]
    cfg = _oracle_cfg(wave, 32, 0.5, 0.02, True)
    ox0, ov0, om = orc.pm_adjoint(pos, vel, m, steps, gxT, None, fixed_e=e, **cfg)
    errs = (_rel(gx0, ox0), _rel(gv0, ov0), _rel(gm, om))
    print("fp32 vs fp64 reference: x0 %.2e v0 %.2e m %.2e" % errs)
    assert max(errs) < 1e-4


def test_pm_adjoint_checkpoint_interval_graphs_and_state(rrg):
    n, steps = 400, 11
    pos, vel, m = ics.exp_disk(n, seed=8)
    rng = np.random.default_rng(2)
    gx, gv = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    ref = None
    for every, graphs in ((1, True), (4, True), (11, True), (3, False)):
        st, g, after = _run(rrg, pos, vel, m, steps, every, gx, gv, graphs=graphs)
        # the live state after the sweep is the recorded run's end state
        assert np.array_equal(st["pos"], after["pos"]) and np.array_equal(st["vel"], after["vel"])
        assert np.array_equal(st["acc"], after["acc"])
        if ref is None:
            ref = g
        else:
            for a, b in zip(ref, g):
                assert _rel(a, b) < 1e-12


def test_pm_adjoint_with_steady_check_matches_oracle(rrg, orc):
    n, steps = 300, 3
    pos, vel, m = ics.exp_disk(n, seed=6)
    wave = _wave(setup_iterations=400, iterations=60, tolerance=2e-3, steady_check_step=5, steady_hold_steps=2)
    rng = np.random.default_rng(11)
    gx = rng.standard_normal((n, 3))
    st, (gx0, gv0, gm), _ = _run(rrg, pos, vel, m, steps, 2, gx, None, wave=wave)
    cfg = _oracle_cfg(wave, 16, 1.0, 0.05, True)
    assert orc.pm_run(pos, vel, m, steps, fixed_e=st["fixed_e"], **cfg)["iterations"][0] < 400
    ox0, ov0, om = orc.pm_adjoint(pos, vel, m, steps, gx, None, fixed_e=st["fixed_e"], **cfg)
    assert _rel(gx0, ox0) < 1e-9 and _rel(gv0, ov0) < 1e-9 and _rel(gm, om) < 1e-9


def test_pm_adjoint_begin_requires_fresh_precalc(rrg):
    pos, vel, m = ics.exp_disk(200, seed=1)
    p = rrg.make_params(dt=0.05, grid=dict(nx=16, dx=1.0), wave=_wave())
    with rrg.Simulation(p, pos, vel, m) as sim:
        sim.precalc()
        sim.calc_step(1)
        with pytest.raises(rrg.RRGError, match="right after rrg_precalc"):
            sim.adjoint_begin(2)
        with pytest.raises(rrg.RRGError, match="without rrg_adjoint_begin"):
            sim.adjoint(np.ones((200, 3)))


def test_torch_pm_run(rrg, orc):
    torch = pytest.importorskip("torch")
    from rrgalaxsim_amd.adjoint import pm_run
    n, steps = 600, 6
    pos, vel, m = ics.exp_disk(n, seed=9)
    wave = _wave()
    p = rrg.make_params(dt=0.05, grid=dict(nx=16, dx=1.0), wave=wave)
    x0 = torch.tensor(pos, requires_grad=True)
    v0 = torch.tensor(vel, requires_grad=True)
    mt = torch.tensor(m, requires_grad=True)
    target = torch.tensor(pos + 0.1)
    xT, vT = pm_run(x0, v0, mt, p, steps, 4)
    e = 61 - math.frexp(float(np.sum(np.abs(m))))[1]  # the library's exponent rule
    loss = ((xT - target) ** 2).mean() + 0.5 * (vT ** 2).sum()
    loss.backward()
    gxT = (2 * (xT - target) / xT.numel()).detach().numpy()
    gvT = vT.detach().numpy()
    cfg = _oracle_cfg(wave, 16, 1.0, 0.05, True)
    ox0, ov0, om = orc.pm_adjoint(pos, vel, m, steps, gxT, gvT, fixed_e=e, **cfg)
    assert _rel(x0.grad.numpy(), ox0) < 1e-9 and _rel(v0.grad.numpy(), ov0) < 1e-9
    assert _rel(mt.grad.numpy(), om) < 1e-9


@pytest.mark.parametrize("real", ["fp64", "fp32"])
def test_pm_adjoint_is_bitwise_reproducible(rrg, real):
    """Two independent runs of the same recorded PM run give bitwise equal
    gradients: the headline-like disk on 64^3 (long runs of equal stencils
    and many atomics per cell)."""
    n, steps = 50_000, 6
    pos, vel, m = ics.exp_disk(n, seed=2)
    rng = np.random.default_rng(6)
    gx, gv = rng.standard_normal((n, 3)) / n, rng.standard_normal((n, 3)) / n
    wave = _wave(setup_iterations=40, iterations=20)
    outs = [_run(rrg, pos, vel, m, steps, 3, gx, gv, real=real, nx=64, dx=0.25, dt=5e-3, wave=wave)[1]
            for _ in range(2)]
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    assert np.max(np.abs(outs[0][2])) > 0
