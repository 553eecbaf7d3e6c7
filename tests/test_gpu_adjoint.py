"""Adjoint of the direct-sum run on the GPU vs the FP64 CPU reverse-mode
gradient of the oracle (itself pinned by finite differences in
tests/test_oracle_nbody.py).  Bars: FP64 kinds agree to 1e-9 relative; the
FP32 kinds' gradients within 1e-4 of the FP64 reference (the target
bound), measured as max |g - g_ref| / max |g_ref| per output.
"""
import numpy as np
import pytest

from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu


def _case(n=512, seed=3):
    pos, vel, m = ics.plummer(n, seed=seed)
    rng = np.random.default_rng(seed)
    m = m * np.exp(0.05 * rng.standard_normal(n))
    s2 = np.full(n, 1e-2 ** 2 * 25)
    return pos, vel, m, s2


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("force,tol", [("fp64", 1e-9), ("fp64_exact", 1e-9), ("mixed", 1e-4), ("fp32", 1e-4)])
def test_adjoint_matches_oracle(rrg, orc, force, tol):
    pos, vel, m, s2 = _case()
    dt, T = 1e-3 * 5, 30
    target = pos + 0.01
    p = rrg.make_params(dt=dt, force=force)
    with rrg.Simulation(p, pos, vel, m, s2) as sim:
        sim.precalc()
        sim.adjoint_begin(7)  # uneven last segment
        sim.calc_step(T)
        xT = sim.state()["pos"]
        gxT = 2 * (xT - target) / xT.size
        gx0, gv0, gm = sim.adjoint(gxT)
    ox0, ov0, om = orc.nbody_adjoint(pos, vel, m, s2, dt, T, gxT)
    print(f"{force}: rel err x0 {_rel(gx0, ox0):.2e} v0 {_rel(gv0, ov0):.2e} m {_rel(gm, om):.2e}")
    assert _rel(gx0, ox0) < tol and _rel(gv0, ov0) < tol and _rel(gm, om) < tol


def test_adjoint_independent_of_checkpoint_interval(rrg):
    pos, vel, m, s2 = _case(300)
    out = []
    for k in (1, 4, 25):
        p = rrg.make_params(dt=2e-3)
        with rrg.Simulation(p, pos, vel, m, s2) as sim:
            sim.precalc()
            sim.adjoint_begin(k)
            sim.calc_step(25)
            out.append(sim.adjoint(np.ones((300, 3)), np.ones((300, 3))))
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert np.array_equal(a, b)


def test_adjoint_velocity_seed_and_state_preserved(rrg, orc):
    """dL/dv_T seeds the sweep too, and the live state is untouched."""
    pos, vel, m, s2 = _case(256)
    p = rrg.make_params(dt=4e-3)
    with rrg.Simulation(p, pos, vel, m, s2) as sim:
        sim.precalc()
        sim.adjoint_begin(5)
        sim.calc_step(12)
        before = sim.state()
        gv = np.random.default_rng(1).standard_normal((256, 3))
        gx0, gv0, gm = sim.adjoint(None, gv)
        after = sim.state()
        assert np.array_equal(before["pos"], after["pos"]) and np.array_equal(before["vel"], after["vel"])
    # oracle: L = sum(gv . v_T) -> seed gv_T = gv
    ox0, ov0, om = orc.nbody_adjoint(pos, vel, m, s2, 4e-3, 12, np.zeros((256, 3)), gv)
    assert _rel(gx0, ox0) < 1e-9 and _rel(gv0, ov0) < 1e-9 and _rel(gm, om) < 1e-9


def test_torch_autograd_function(rrg, orc):
    torch = pytest.importorskip("torch")
    from rrgalaxsim_amd.adjoint import direct_sum_run

    pos, vel, m, s2 = _case(400)
    target = torch.tensor(pos + 0.02)
    x0 = torch.tensor(pos, requires_grad=True)
    v0 = torch.tensor(vel, requires_grad=True)
    mt = torch.tensor(m, requires_grad=True)
    xT, vT = direct_sum_run(x0, v0, mt, torch.tensor(s2), 3e-3, 20, 6, "fp64")
    loss = ((xT - target) ** 2).mean()
    loss.backward()
    gx = 2 * (xT.detach().numpy() - target.numpy()) / xT.numel()
    ox0, ov0, om = orc.nbody_adjoint(pos, vel, m, s2, 3e-3, 20, gx)
    assert _rel(x0.grad.numpy(), ox0) < 1e-9
    assert _rel(v0.grad.numpy(), ov0) < 1e-9
    assert _rel(mt.grad.numpy(), om) < 1e-9


def test_adjoint_without_begin_is_an_error(rrg):
    pos, vel, m, s2 = _case(64)
    with rrg.Simulation(rrg.make_params(dt=1e-3), pos, vel, m, s2) as sim:
        sim.precalc()
        with pytest.raises(rrg.RRGError) as ei:
            sim.adjoint(np.ones((64, 3)))
        assert ei.value.code == rrg._lib.RRG_ERR_STATE
