"""torch.autograd wrappers of the HIP discrete adjoint (no reference
counterpart: the reference computes no derivatives).

  vortex_state(params, solver, t=0)   U0 of the type 0 vortex (Element.cpp:119-179)
                                      as a function of params = (epsilon, x0, y0);
                                      backward through vt_vortex_tangent
  rk4_solve(U0, solver, dt, nsteps)   U_T after nsteps RK4 steps (Timestepping.cpp:85-305)
                                      on the device; backward = the checkpointed
                                      reverse sweep (vt_adjoint_forward/backward);
                                      several calls on one solver may share a
                                      graph: a node whose recording was replaced
                                      records its run again before its sweep
  mass_l2(U, solver, rho_target)      J = sum_e detJ_e drho^T M_ref drho in torch

Tensors are float64 of shape [n][Np][4] in the reference's element-major
order; the solver's device state is set from U0 and left at U_T.
"""
import numpy as np
import torch


class VortexState(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, solver, t):
        eps, x0, y0 = (float(v) for v in params.detach().cpu().tolist())
        ctx.solver, ctx.t, ctx.p = solver, t, (eps, x0, y0)
        return torch.from_numpy(solver.exact_vortex(t, eps, x0, y0)).to(params.device)

    @staticmethod
    def backward(ctx, g):
        T = ctx.solver.vortex_tangent(ctx.t, *ctx.p)
        gp = np.tensordot(T, g.detach().cpu().double().numpy(), axes=([1, 2, 3], [0, 1, 2]))
        return torch.from_numpy(gp).to(g.device), None, None


class RK4Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, solver, dt, nsteps, every):
        u0 = U0.detach().cpu().double().numpy().copy()
        solver.set_state(u0)
        solver.adjoint_forward(dt, nsteps, every)
        ctx.solver, ctx.u0, ctx.run = solver, u0, (dt, nsteps, every)
        ctx.rec = solver.adjoint_recording()
        return torch.from_numpy(solver.state()).to(U0.device)

    @staticmethod
    def backward(ctx, g):
        s = ctx.solver
        if s.adjoint_recording() != ctx.rec:
            # Another run was recorded on this solver since (a second rk4_solve
            # in the same graph, or an adjoint call): record this node's run
            # again from its saved U0, leaving the solver's state and time as
            # they were.
            cur, t = s.state(), s.time
            s.set_state(ctx.u0)
            s.adjoint_forward(*ctx.run)
            ctx.rec = s.adjoint_recording()
            s.set_state(cur)
            s.time = t
        lam0 = s.adjoint_backward(g.detach().cpu().double().numpy())
        return torch.from_numpy(lam0).to(g.device), None, None, None, None


def vortex_state(params, solver, t=0.0):
    return VortexState.apply(params, solver, t)


def rk4_solve(U0, solver, dt, nsteps, every=32):
    return RK4Solve.apply(U0, solver, dt, nsteps, every)


def mass_l2(U, solver, rho_target):
    """J = sum_e detJ_e (rho_e - rho_target_e)^T M_ref (rho_e - rho_target_e)."""
    from . import reference_operators
    M, _, _ = reference_operators(solver.p, solver.order)
    v = solver.mesh.element_coordinates
    det = (v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) - (v[:, 1, 1] - v[:, 0, 1]) * (v[:, 2, 0] - v[:, 0, 0])
    d = U[:, :, 0] - torch.as_tensor(rho_target, dtype=U.dtype, device=U.device)
    Mt = torch.as_tensor(M, dtype=U.dtype, device=U.device)
    return (torch.as_tensor(det, dtype=U.dtype, device=U.device) * ((d @ Mt) * d).sum(1)).sum()
