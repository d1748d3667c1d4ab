"""torch.autograd entry points of the variable-density pressure solve and of
the whole pressure projection.

``p = pressure_solve(ctx, log_rho, rhs, dt)`` solves ``A(rho) p = rhs`` with
the SP operator of src/module_simulation.f90:2483-2533 (face coefficients
``dt*dy/(rho_f*dx)``, ``rho_f`` the arithmetic mean of the two cell densities
as in src/module_interface_vof.f90:574-590), for a cell density given by its
logarithm on the interior ``(ny, nx)``.  Halo densities follow the reference's
halo fill (periodic wrap, zero-gradient copy otherwise).

The backward pass is the hand-written adjoint of libcovof_hip
(``covof_poisson_adjoint_d``): one transpose solve (A is symmetric) plus the
adjoint of the face-density averaging and of the coefficient formula; it
returns dJ/d(log rho) and dJ/d(rhs).

``u, v, p = projection(ctx, log_rho, u_star, v_star, dt)`` is the whole
projection of the standard projection (src/module_simulation.f90:2477-2654):
b = -div(u*), the solve, p = x - mean(x), and the density-weighted gradient
correction of u*, v*.  Its backward (``covof_projection_adjoint_d``) returns
dJ/d(log rho), dJ/du* and dJ/dv* for any objective of (u, v, p).
"""
from __future__ import annotations

import torch

from . import _lib as L


def _rho_fields(ctx, log_rho: torch.Tensor):
    nx, ny = ctx.nx, ctx.ny
    rho = torch.empty((ny + 2, nx + 2), dtype=torch.float64, device=log_rho.device)
    rho[1:ny + 1, 1:nx + 1] = torch.exp(log_rho)
    perx = ctx.bc[0] == L.BC_PERIODIC or ctx.bc[1] == L.BC_PERIODIC
    pery = ctx.bc[2] == L.BC_PERIODIC or ctx.bc[3] == L.BC_PERIODIC
    if perx:
        rho[1:ny + 1, 0] = rho[1:ny + 1, nx]
        rho[1:ny + 1, nx + 1] = rho[1:ny + 1, 1]
    else:
        rho[1:ny + 1, 0] = rho[1:ny + 1, 1]
        rho[1:ny + 1, nx + 1] = rho[1:ny + 1, nx]
    if pery:
        rho[0, :] = rho[ny, :]
        rho[ny + 1, :] = rho[1, :]
    else:
        rho[0, :] = rho[1, :]
        rho[ny + 1, :] = rho[ny, :]
    rho_u = torch.ones_like(rho)
    rho_v = torch.ones_like(rho)
    rho_u[1:ny + 1, 1:nx + 2] = 0.5 * (rho[1:ny + 1, 0:nx + 1] + rho[1:ny + 1, 1:nx + 2])
    rho_v[1:ny + 2, 1:nx + 1] = 0.5 * (rho[0:ny + 1, 1:nx + 1] + rho[1:ny + 2, 1:nx + 1])
    return rho, rho_u, rho_v


class PressureSolve(torch.autograd.Function):
    @staticmethod
    def forward(fctx, ctx, log_rho, rhs, dt, solver):
        rho, rho_u, rho_v = _rho_fields(ctx, log_rho.detach())
        zero = ctx.new_field()
        ctx.poisson_assemble(zero, zero, rho_u, rho_v, dt)
        x = ctx.new_work()
        res = ctx.poisson_solve(rhs.detach().contiguous(), x, **solver)
        if not res.converged:
            raise L.CovofError(res.status, "pressure_solve forward",
                               f"iters={res.iters} relres={res.relres:.3e}")
        fctx.ctx = ctx
        fctx.dt = dt
        fctx.solver = solver
        fctx.save_for_backward(rho, rho_u, rho_v, x)
        fctx.forward_iters = res.iters
        return x

    @staticmethod
    def backward(fctx, grad_p):
        ctx = fctx.ctx
        rho, rho_u, rho_v, x = fctx.saved_tensors
        zero = ctx.new_field()
        ctx.poisson_assemble(zero, zero, rho_u, rho_v, fctx.dt)  # the operator of this forward
        g_b, g_lr, res = ctx.poisson_adjoint(rho, fctx.dt, x, grad_p.contiguous(), **fctx.solver)
        if not res.converged:
            raise L.CovofError(res.status, "pressure_solve backward",
                               f"iters={res.iters} relres={res.relres:.3e}")
        return None, g_lr, g_b, None, None


class Projection(torch.autograd.Function):
    @staticmethod
    def forward(fctx, ctx, log_rho, u_star, v_star, dt, solver):
        rho, rho_u, rho_v = _rho_fields(ctx, log_rho.detach())
        us, vs = u_star.detach().contiguous(), v_star.detach().contiguous()
        b = ctx.poisson_assemble(us, vs, rho_u, rho_v, dt)
        x = ctx.new_work()
        res = ctx.poisson_solve(b, x, **solver)
        if not res.converged:
            raise L.CovofError(res.status, "projection forward", f"iters={res.iters} relres={res.relres:.3e}")
        p = ctx.new_field()
        ctx.sp_pressure(x, p)
        u, v = us.clone(), vs.clone()
        ctx.correct_velocity(p, rho_u, rho_v, dt, u, v)
        fctx.ctx, fctx.dt, fctx.solver = ctx, dt, solver
        fctx.save_for_backward(rho, rho_u, rho_v, us, vs, x)
        return u, v, p

    @staticmethod
    def backward(fctx, gu, gv, gp):
        ctx = fctx.ctx
        rho, rho_u, rho_v, us, vs, x = fctx.saved_tensors
        ctx.poisson_assemble(us, vs, rho_u, rho_v, fctx.dt)  # the operator of this forward
        sol = {k: v for k, v in fctx.solver.items() if k != "linear_solver"}
        zero = lambda g: g.contiguous() if g is not None else ctx.new_field()
        du, dv, dlr, res = ctx.projection_adjoint(rho, rho_u, rho_v, fctx.dt, x, zero(gu), zero(gv),
                                                  gp.contiguous() if gp is not None else None, **sol)
        if not res.converged:
            raise L.CovofError(res.status, "projection backward", f"iters={res.iters} relres={res.relres:.3e}")
        return None, dlr, du, dv, None, None


def projection(ctx, log_rho: torch.Tensor, u_star: torch.Tensor, v_star: torch.Tensor, dt: float, *,
               precond=L.PRECOND_JACOBI, tol=1e-12, max_iter=200000, precision=L.PRECISION_FP64):
    """Differentiable projection (u, v, p) of the face velocities u*, v* for
    the cell density exp(log_rho) (interior (ny, nx); halo densities by the
    reference's halo fill, face densities their arithmetic means)."""
    solver = dict(linear_solver=L.LINEAR_CG, precond=precond, tol=tol, max_iter=max_iter, precision=precision)
    return Projection.apply(ctx, log_rho, u_star, v_star, dt, solver)


def pressure_solve(ctx, log_rho: torch.Tensor, rhs: torch.Tensor, dt: float = 1.0, *,
                   linear_solver=L.LINEAR_CG, precond=L.PRECOND_JACOBI, tol=1e-12,
                   max_iter=200000, precision=L.PRECISION_FP64) -> torch.Tensor:
    """Differentiable ``p = A(exp(log_rho))^-1 rhs`` on the interior grid."""
    solver = dict(linear_solver=linear_solver, precond=precond, tol=tol, max_iter=max_iter,
                  precision=precision)
    return PressureSolve.apply(ctx, log_rho, rhs, dt, solver)
