"""torch.autograd bindings of the discrete adjoint (include/dg.h, dg_adjoint_*).

DGObjective: J(u0, eps) = objective(u_T) after nsteps fixed-dt steps; the
backward pass returns dJ/du0 and (Maxwell TE) dJ/deps from one checkpointed
adjoint sweep on the GPU.

DGRhs: k = dt * L(u) for one evaluation of the right-hand side, with the
vector-Jacobian product (dk/du)^T g as its backward pass, so torch code can
compose its own time integrators on top of the HIP operator.

Tensors are float64; host (CPU) tensors are copied through the C ABI's host
path, CUDA tensors are passed as device pointers.
"""
import numpy as np
import torch

from . import _lib


def _np(t):
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


class DGObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, eps, solver, nsteps, dt, every):
        solver.set_coeffs(_np(u0).reshape(solver.shape))
        want_eps = eps is not None and solver.physics == _lib.MAXWELL_TE
        if want_eps:
            solver.set_eps(_np(eps))
        J, gu, ge = solver.adjoint_solve(int(nsteps), float(dt), every=int(every),
                                         want_eps=want_eps)
        ctx.save_for_backward(torch.from_numpy(gu).to(u0.device),
                              torch.from_numpy(ge).to(u0.device) if want_eps else
                              torch.zeros(0, dtype=torch.float64, device=u0.device))
        ctx.want_eps = want_eps
        return torch.tensor(J, dtype=torch.float64, device=u0.device)

    @staticmethod
    def backward(ctx, gJ):
        gu, ge = ctx.saved_tensors
        geps = gJ * ge if ctx.want_eps else None
        return gJ * gu, geps, None, None, None, None


def objective(solver, u0, nsteps, dt, eps=None, every=0):
    """J(u0[, eps]) as a differentiable torch scalar."""
    return DGObjective.apply(u0, eps, solver, nsteps, dt, every)


class DGRhs(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, solver, t, dt):
        k = solver.eval_rhs(_np(u).reshape(solver.shape), t=float(t), dt=float(dt))
        ctx.save_for_backward(u.detach().clone())
        ctx.solver, ctx.t, ctx.dt = solver, float(t), float(dt)
        return torch.from_numpy(np.ascontiguousarray(k)).to(u.device)

    @staticmethod
    def backward(ctx, gk):
        (u,) = ctx.saved_tensors
        gu = ctx.solver.adjoint_rhs(_np(u).reshape(ctx.solver.shape),
                                    _np(gk).reshape(ctx.solver.shape), t=ctx.t, dt=ctx.dt)
        return torch.from_numpy(gu).to(gk.device), None, None, None


def rhs(solver, u, t=0.0, dt=1.0):
    """k = dt * L(u) as a differentiable torch op."""
    return DGRhs.apply(u, solver, t, dt)
