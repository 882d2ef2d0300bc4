"""torch.autograd binding of the discrete adjoint (astr_adjoint_record and
astr_adjoint_reverse).

EnstrophyAfterSteps.apply(q0, prim0, solver, nsteps, dt, ckpt) returns the
enstrophy J after nsteps RK3 steps of the plain solver from the state
(q0, prim0); its backward is the hand-written GPU adjoint of those steps.
q0 and prim0 are CPU float64 tensors in the haloed host layout of
AstrSolver.upload_state, q0 (ni, nj, nk, 5) and prim0 (ni, nj, nk, 6) =
(rho, u, v, w, p, T); the solver is the AstrSolver that owns the box (its
geometry uploaded).  velocity_state builds that state from a velocity field
as state_from_velocity does, in torch, so that dJ/du0 follows by the chain
rule:

    u0 = torch.tensor(u, requires_grad=True)
    q0, prim0 = velocity_state(params, u0)
    J = EnstrophyAfterSteps.apply(q0, prim0, solver, 40, 0.002, 8)
    J.backward()          # u0.grad = dJ/du0
"""
from __future__ import annotations

import numpy as np
import torch

from .solver import HM, AstrParams, AstrSolver


def _upload(solver: AstrSolver, q0: torch.Tensor, prim0: torch.Tensor) -> None:
    q = np.asfortranarray(q0.detach().cpu().numpy())
    pr = prim0.detach().cpu().numpy()
    solver.upload_state(q, np.asfortranarray(pr[..., 0]), np.asfortranarray(pr[..., 1:4]),
                        np.asfortranarray(pr[..., 4]), np.asfortranarray(pr[..., 5]))


class EnstrophyAfterSteps(torch.autograd.Function):
    """Forward: nsteps plain RK3 steps on the GPU (astr_adjoint_record: the
    state is kept every ckpt steps and the last segment's stages are stored
    on the device as they run), then J = the enstrophy statistic of the
    final state after a qswap.  Backward: astr_adjoint_reverse of that
    recorded trajectory (the segments before the last are recomputed from
    their kept states).  The solver keeps one recorded trajectory: a second
    forward on the same solver before this backward replaces it, and the
    backward then raises."""

    @staticmethod
    def forward(ctx, q0, prim0, solver, nsteps, dt, ckpt):
        _upload(solver, q0, prim0)
        J = solver.adjoint_record(int(nsteps), float(dt), int(ckpt))
        solver._adjoint_token = token = object()
        ctx.solver, ctx.token = solver, token
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        if getattr(ctx.solver, "_adjoint_token", None) is not ctx.token:
            raise RuntimeError("the solver recorded another trajectory after this forward")
        qb, pb, _ = ctx.solver.adjoint_reverse()
        g = gJ.to(torch.float64)
        return (g * torch.from_numpy(np.ascontiguousarray(qb)),
                g * torch.from_numpy(np.ascontiguousarray(pb)), None, None, None, None)


def velocity_state(p: AstrParams, u0: torch.Tensor):
    """(q0, prim0) of state_from_velocity in torch: u0 (3, n0, n1, n2) at the
    unique nodes; rho = 1, T = 1, p = 1/(gamma Ma^2), q by fvar2q's
    temperature form; node n copies node 0; halos zero."""
    n = (p.im, p.jm, p.km)
    const1 = 1.0 / (p.gamma * (p.gamma - 1.0) * (p.mach * p.mach))
    const2 = p.gamma * (p.mach * p.mach)
    idx = [torch.arange(k + 1) % k for k in n]
    u = u0[:, idx[0]][:, :, idx[1]][:, :, :, idx[2]].to(torch.float64)
    rho = torch.ones_like(u[0])
    prs = torch.full_like(u[0], 1.0 / const2)
    tmp = prs / rho * const2
    shape = tuple(k + 1 + 2 * HM for k in n)
    I = tuple(slice(HM, HM + k + 1) for k in n)
    q0 = torch.zeros(shape + (5,), dtype=torch.float64)
    pr0 = torch.zeros(shape + (6,), dtype=torch.float64)
    q0 = q0.clone()
    pr0 = pr0.clone()
    q0[I] = torch.stack([rho, rho * u[0], rho * u[1], rho * u[2],
                         rho * (tmp * const1 + 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]))],
                        dim=-1)
    pr0[I] = torch.stack([rho, u[0], u[1], u[2], prs, tmp], dim=-1)
    return q0, pr0
