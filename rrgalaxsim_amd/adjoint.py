"""torch.autograd.Function bindings of the library's adjoints.

``DirectSumRun.apply(x0, v0, m, soft2, dt, nsteps, checkpoint_every, force)``
runs ``nsteps`` velocity-Verlet steps of the direct-sum model (the reference's
predict_step_ / acceleration_kernel_blocked_ / correct_step_, nbody.cu:19-181)
on the GPU and returns (x_T, v_T).  Its backward pass calls the library's
hand-written adjoint (rrg_adjoint: checkpoint-every-K recomputation and the
transposed Verlet step with one pair-force vector-Jacobian product per step)
and returns dL/dx0, dL/dv0 and dL/dm.  Tensors cross the C ABI as float64
host arrays; the solver state lives in the library's device memory.

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


class DirectSumRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, m, soft2, dt: float, nsteps: int, checkpoint_every: int = 20, force: str = "mixed"):
        p = S.make_params(dt=dt, force=force)
        sim = S.Simulation(p, _np(x0), _np(v0), _np(m), _np(soft2))
        sim.precalc()
        sim.adjoint_begin(checkpoint_every)
        sim.calc_step(nsteps)
        st = sim.state()
        ctx.sim = sim
        ctx.meta = (x0.device, x0.dtype, v0.dtype, m.dtype)
        dev, dx, dv, _ = ctx.meta
        return torch.from_numpy(st["pos"]).to(dev, dx), torch.from_numpy(st["vel"]).to(dev, dv)

    @staticmethod
    def backward(ctx, gx_T, gv_T):
        dev, dx, dv, dm = ctx.meta
        gx = None if gx_T is None else _np(gx_T)
        gv = None if gv_T is None else _np(gv_T)
        gx0, gv0, gm = ctx.sim.adjoint(gx, gv)
        ctx.sim.close()
        ctx.sim = None
        return (torch.from_numpy(gx0).to(dev, dx), torch.from_numpy(gv0).to(dev, dv),
                torch.from_numpy(gm).to(dev, dm), None, None, None, None, None)


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
        st = sim.state()
        ctx.sim = sim
        ctx.meta = (x0.device, x0.dtype, v0.dtype, m.dtype)
        dev, dx, dv, _ = ctx.meta
        return torch.from_numpy(st["pos"]).to(dev, dx), torch.from_numpy(st["vel"]).to(dev, dv)

    @staticmethod
    def backward(ctx, gx_T, gv_T):
        dev, dx, dv, dm = ctx.meta
        gx = None if gx_T is None else _np(gx_T)
        gv = None if gv_T is None else _np(gv_T)
        gx0, gv0, gm = ctx.sim.adjoint(gx, gv)
        ctx.sim.close()
        ctx.sim = None
        return (torch.from_numpy(gx0).to(dev, dx), torch.from_numpy(gv0).to(dev, dv),
                torch.from_numpy(gm).to(dev, dm), None, None, None)


def pm_run(x0, v0, m, params, nsteps, checkpoint_every=20):
    """Functional form of PMRun.apply."""
    return PMRun.apply(x0, v0, m, params, nsteps, checkpoint_every)
