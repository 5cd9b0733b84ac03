"""torch.autograd bindings of the discrete adjoint (include/dg.h, dg_adjoint_*).

DGObjective: J(u0, eps) = objective(u_T) after nsteps fixed-dt steps; the
backward pass returns dJ/du0 and (Maxwell TE) dJ/deps from one checkpointed
adjoint sweep on the GPU.

DGRhs: k = dt * L(u) for one evaluation of the right-hand side, with the
vector-Jacobian product (dk/du)^T g as its backward pass, so torch code can
compose its own time integrators on top of the HIP operator.

Tensors are float64 with N * n_p * K elements ([N, n_p, K] in the
reference's SoA order).  CUDA tensors on the solver's device cross the C ABI
as device pointers (DG_DEVICE): the state, k, the gradients and the VJP stay
in device memory with no host copy.  CPU tensors take the host path
(DG_HOST).  The permittivity eps ([K]) always crosses as host data.
"""
import torch

from . import _lib


def _state(solver, t):
    return t.detach().reshape(solver.shape)


class DGObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, eps, solver, nsteps, dt, every):
        solver.set_coeffs(_state(solver, u0))
        want_eps = eps is not None and solver.physics == _lib.MAXWELL_TE
        if want_eps:
            solver.set_eps(eps.detach())
        dev = u0.device if u0.device.type == "cuda" else None
        J, gu, ge = solver.adjoint_solve(int(nsteps), float(dt), every=int(every),
                                         want_eps=want_eps, device=dev)
        if dev is None:
            gu = torch.from_numpy(gu)
            ge = torch.from_numpy(ge) if want_eps else None
        ctx.save_for_backward(gu.reshape(u0.shape),
                              ge.to(eps.device) if want_eps else
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
        us = _state(solver, u).contiguous()
        k = solver.eval_rhs(us, t=float(t), dt=float(dt))
        if not torch.is_tensor(k):
            k = torch.from_numpy(k)
        ctx.save_for_backward(us.clone())
        ctx.solver, ctx.t, ctx.dt = solver, float(t), float(dt)
        return k.reshape(u.shape)

    @staticmethod
    def backward(ctx, gk):
        (u,) = ctx.saved_tensors
        g = gk.detach().reshape(ctx.solver.shape).to(u.device, torch.float64).contiguous()
        gu = ctx.solver.adjoint_rhs(u, g, t=ctx.t, dt=ctx.dt)
        if not torch.is_tensor(gu):
            gu = torch.from_numpy(gu)
        return gu.reshape(gk.shape), None, None, None


def rhs(solver, u, t=0.0, dt=1.0):
    """k = dt * L(u) as a differentiable torch op."""
    return DGRhs.apply(u, solver, t, dt)
