"""torch.autograd.Function bindings of the library's adjoints.

``DirectSumRun.apply(x0, v0, m, soft2, dt, nsteps, checkpoint_every, force)``
runs ``nsteps`` velocity-Verlet steps of the direct-sum model (the reference's
predict_step_ / acceleration_kernel_blocked_ / correct_step_, nbody.cu:19-181)
on the GPU and returns (x_T, v_T).  Its backward pass calls the library's
hand-written adjoint (rrg_adjoint: checkpoint-every-K recomputation and the
transposed Verlet step with one pair-force vector-Jacobian product per step)
and returns dL/dx0, dL/dv0 and dL/dm.  The solver state lives in the
library's device memory.  CPU tensors cross the C ABI as float64 host arrays;
for tensors on the GPU the outputs, seeds and gradients move device to device
(rrg_get_state_dev / rrg_adjoint_dev, ordered on torch's current stream) and
only the initial state is copied through the host once.

``PMRun.apply(x0, v0, m, params, nsteps, checkpoint_every)`` does the same for
the particle-mesh model (params with a grid and wave section, e.g. from
``solver.make_params(grid=..., wave=...)``): precalc, the synchronised a_0,
``nsteps`` PM steps (CIC deposit, damped-wave iterations, -grad phi,
trilinear interpolation; rrgsim_integration_rect.cuh:53-112), and in backward
the reversed grid builds (interpolation^T = deposit, CIC^T = interpolation,
the transposed stencil recurrence) with checkpointed recomputation.
"""
from __future__ import annotations

import numpy as np
import torch

from . import solver as S


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _final_state(sim, x0, v0):
    """(x_T, v_T) in x0's and v0's device and dtype."""
    if x0.is_cuda:
        x = torch.empty((sim.n, 3), dtype=torch.float64, device=x0.device)
        v = torch.empty((sim.n, 3), dtype=torch.float64, device=x0.device)
        sim.get_state_dev(pos=x.data_ptr(), vel=v.data_ptr(), stream=_stream(x0))
        return x.to(x0.dtype), v.to(v0.dtype)
    st = sim.state()
    return torch.from_numpy(st["pos"]).to(x0.device, x0.dtype), torch.from_numpy(st["vel"]).to(v0.device, v0.dtype)


def _gradients(sim, gx_T, gv_T, meta):
    """(dL/dx0, dL/dv0, dL/dm) for the seeds, in the inputs' devices and dtypes."""
    dev, dx, dv, dm = meta
    if dev.type == "cuda":
        seeds = [None if g is None else g.detach().to(dev, torch.float64).contiguous() for g in (gx_T, gv_T)]
        gx0 = torch.empty((sim.n, 3), dtype=torch.float64, device=dev)
        gv0 = torch.empty((sim.n, 3), dtype=torch.float64, device=dev)
        gm = torch.empty((sim.n,), dtype=torch.float64, device=dev)
        sim.adjoint_dev(gx_T=0 if seeds[0] is None else seeds[0].data_ptr(),
                        gv_T=0 if seeds[1] is None else seeds[1].data_ptr(),
                        gx_0=gx0.data_ptr(), gv_0=gv0.data_ptr(), gm=gm.data_ptr(),
                        stream=torch.cuda.current_stream(dev).cuda_stream)
        return gx0.to(dx), gv0.to(dv), gm.to(dm)
    gx = None if gx_T is None else _np(gx_T)
    gv = None if gv_T is None else _np(gv_T)
    gx0, gv0, gm = sim.adjoint(gx, gv)
    return torch.from_numpy(gx0).to(dev, dx), torch.from_numpy(gv0).to(dev, dv), torch.from_numpy(gm).to(dev, dm)


class DirectSumRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, m, soft2, dt: float, nsteps: int, checkpoint_every: int = 20, force: str = "mixed"):
        device = (x0.device.index or 0) if x0.is_cuda else 0
        p = S.make_params(dt=dt, force=force, device=device)
        sim = S.Simulation(p, _np(x0), _np(v0), _np(m), _np(soft2))
        sim.precalc()
        sim.adjoint_begin(checkpoint_every)
        sim.calc_step(nsteps)
        ctx.sim = sim
        ctx.meta = (x0.device, x0.dtype, v0.dtype, m.dtype)
        return _final_state(sim, x0, v0)

    @staticmethod
    def backward(ctx, gx_T, gv_T):
        gx0, gv0, gm = _gradients(ctx.sim, gx_T, gv_T, ctx.meta)
        ctx.sim.close()
        ctx.sim = None
        return gx0, gv0, gm, None, None, None, None, None


def direct_sum_run(x0, v0, m, soft2, dt, nsteps, checkpoint_every=20, force="mixed"):
    """Functional form of DirectSumRun.apply."""
    return DirectSumRun.apply(x0, v0, m, soft2, dt, nsteps, checkpoint_every, force)


class PMRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, m, params, nsteps: int, checkpoint_every: int = 20):
        if not params.has_grid or not params.has_wave:
            raise ValueError("PMRun: params need a grid and a wave section")
        sim = S.Simulation(params, _np(x0), _np(v0), _np(m))
        sim.precalc()
        sim.adjoint_begin(checkpoint_every)
        sim.calc_step(nsteps)
        ctx.sim = sim
        ctx.meta = (x0.device, x0.dtype, v0.dtype, m.dtype)
        return _final_state(sim, x0, v0)

    @staticmethod
    def backward(ctx, gx_T, gv_T):
        gx0, gv0, gm = _gradients(ctx.sim, gx_T, gv_T, ctx.meta)
        ctx.sim.close()
        ctx.sim = None
        return gx0, gv0, gm, None, None, None


def pm_run(x0, v0, m, params, nsteps, checkpoint_every=20):
    """Functional form of PMRun.apply."""
    return PMRun.apply(x0, v0, m, params, nsteps, checkpoint_every)
