"""Device-array entry points (rrg_get_state_dev, rrg_set_state_dev,
rrg_adjoint_dev): the same values as the host-buffer entry points, in the
caller's particle order also after PM contexts re-sort their particles,
ordered on the caller's stream; host pointers are rejected; the torch binding
gives the same results for GPU and CPU tensors."""
import numpy as np
import pytest
import torch

from rrgalaxsim_amd import adjoint as A
from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu

C = 4574.337022617616


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


def _cur():
    return torch.cuda.current_stream().cuda_stream


def _nbody(rrg, n=3000, force="fp64"):
    pos, vel, m = ics.plummer(n, seed=4)
    return rrg.Simulation(rrg.make_params(dt=1e-3, force=force), pos, vel, m, np.full(n, 1e-4)), pos, vel


def _pm(rrg, n=5000, real="fp32"):
    pos, vel, m = ics.exp_disk(n, seed=6)
    wave = dict(diss_base=0.785, diss_extra=0.0, wave_speed=C, setup_iterations=20, iterations=5, tolerance=0.0)
    p = rrg.make_params(dt=5e-3, grid=dict(nx=32, dx=0.5), wave=wave, real=real, sort_interval=8)
    return rrg.Simulation(p, pos, vel, m), pos, vel


@pytest.mark.parametrize("model", ["nbody", "pm"])
def test_get_state_dev_equals_host(rrg, model):
    sim, pos, vel = _nbody(rrg) if model == "nbody" else _pm(rrg)
    with sim:
        sim.precalc()
        sim.sync_particles_and_grid()
        sim.calc_step(20)  # PM contexts re-sort their particles on the way
        sim.sync_particles_and_grid()
        host = sim.state()
        d = [torch.full((sim.n, 3), np.nan, dtype=torch.float64, device="cuda") for _ in range(3)]
        sim.get_state_dev(pos=d[0].data_ptr(), vel=d[1].data_ptr(), acc=d[2].data_ptr(), stream=_cur())
        torch.cuda.current_stream().synchronize()
        for k, t in zip(("pos", "vel", "acc"), d):
            assert np.array_equal(t.cpu().numpy(), host[k]), k


@pytest.mark.parametrize("model", ["nbody", "pm"])
def test_set_state_dev_equals_host(rrg, model):
    out = []
    for dev in (False, True):
        sim, pos, vel = _nbody(rrg) if model == "nbody" else _pm(rrg)
        with sim:
            sim.precalc()
            sim.calc_step(18)
            newp, newv = pos * 0.9 + 0.01, vel * 1.1
            if dev:
                tp, tv = _t(newp), _t(newv)
                sim.set_state_dev(pos=tp.data_ptr(), vel=tv.data_ptr(), stream=_cur())
            else:
                sim.set_state(pos=newp, vel=newv)
            st0 = sim.state()
            sim.sync_particles_and_grid()
            sim.calc_step(3)
            out.append((st0, sim.state()))
    (a0, a1), (b0, b1) = out
    for k in ("pos", "vel"):
        assert np.array_equal(a0[k], b0[k]) and np.array_equal(a1[k], b1[k]), k


@pytest.mark.parametrize("model", ["nbody", "pm"])
def test_adjoint_dev_equals_host(rrg, model):
    res = []
    for dev in (False, True):
        sim, pos, vel = _nbody(rrg, n=1024) if model == "nbody" else _pm(rrg, n=2000, real="fp64")
        rng = np.random.default_rng(3)
        gx, gv = rng.standard_normal((sim.n, 3)), rng.standard_normal((sim.n, 3))
        with sim:
            sim.precalc()
            sim.adjoint_begin(3)
            sim.calc_step(7)
            if dev:
                tg = [_t(gx), _t(gv)]
                o = [torch.empty((sim.n, 3), dtype=torch.float64, device="cuda") for _ in range(2)]
                om = torch.empty((sim.n,), dtype=torch.float64, device="cuda")
                sim.adjoint_dev(gx_T=tg[0].data_ptr(), gv_T=tg[1].data_ptr(), gx_0=o[0].data_ptr(),
                                gv_0=o[1].data_ptr(), gm=om.data_ptr(), stream=_cur())
                torch.cuda.current_stream().synchronize()
                res.append((o[0].cpu().numpy(), o[1].cpu().numpy(), om.cpu().numpy()))
            else:
                res.append(sim.adjoint(gx, gv))
    for a, b in zip(*res):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def test_host_pointers_are_rejected(rrg):
    sim, pos, vel = _nbody(rrg, n=256)
    with sim:
        sim.precalc()
        host = np.empty((sim.n, 3))
        with pytest.raises(rrg.RRGError, match="not device memory"):
            sim.get_state_dev(pos=host.ctypes.data)


def test_torch_binding_gpu_tensors_match_cpu_tensors(rrg):
    pos, vel, m = ics.plummer(512, seed=5)
    s2 = np.full(512, 1e-4)
    grads = []
    for device in ("cpu", "cuda"):
        x0 = torch.tensor(pos, dtype=torch.float64, device=device, requires_grad=True)
        v0 = torch.tensor(vel, dtype=torch.float64, device=device, requires_grad=True)
        mm = torch.tensor(m, dtype=torch.float64, device=device, requires_grad=True)
        ss = torch.tensor(s2, dtype=torch.float64, device=device)
        xT, vT = A.direct_sum_run(x0, v0, mm, ss, 1e-3, 6, checkpoint_every=2, force="fp64")
        assert xT.device.type == device
        loss = (xT ** 2).sum() + 0.5 * (vT ** 2).sum()
        loss.backward()
        grads.append([t.grad.detach().cpu().numpy() for t in (x0, v0, mm)] + [xT.detach().cpu().numpy()])
    for a, b in zip(*grads):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
