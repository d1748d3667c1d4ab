"""Gradient of the kinetic energy after N two-phase steps with respect to the
initial volume fraction C0, by a checkpointed discrete adjoint (no reference
counterpart).

Forward step n -> n+1, in the order of the reference's time loop for the
standard projection (src/module_simulation.f90:495-559,
``sim_solve_standard_projection`` :2424-2666) with the momentum block reduced
to its BDF2 inertia and gravity (convection, viscosity and surface tension
off):

1. ``advect_vof`` (Weymouth-Yue, PLIC or ELVIRA; pass order from the step
   number; the two-phase corner fix of the halo): C^n -> C^{n+1} with u^n, v^n;
   rho, rho_u, rho_v from C^{n+1} (update_physical_properties);
2. u* = (4 u^n - u^{n-1} + 2 dt g) / 3 (u^n + dt g on the first step), zero
   on the wall faces;
3. the Poisson block: b = -div u*, A(rho) x = b by CG from x = 0 to a tight
   tolerance, p = x - mean(x), and the correction u^{n+1} = u* - dt/rho_u dp/dx.

J = the history row's kinetic energy of (C^N, u^N, v^N)
(src/module_vtk_writer.f90:399-531).

The forward stores, per step, C^n, u^n, v^n and the solve's x (32 bytes per
cell per step: 4.3 GB for 16 steps at 4096 x 2048).  The backward walks the
steps in reverse and chains the library's adjoint entry points: the
projection adjoint (the transpose solve on the re-assembled operator), the
density update's, the predictor's transpose and the Weymouth-Yue step's.
Each step's intermediate fields (props, u*, the operator) are recomputed
from the stored states.  Every kernel runs in libcovof_hip; torch holds the
arrays.
"""
from __future__ import annotations

import torch

from . import _lib as L
from .context import Context


class KineticEnergyAdjoint:
    """Forward run with per-step checkpoints, then the reverse sweep.

    ``ctx``: a context whose sides are all walls (DIRICHLET / FREESLIP) or
    all periodic.  ``tol``: relative residual of the forward and transpose
    CG solves (Jacobi-preconditioned unless ``precond`` says otherwise); the
    gradient is that of the exactly solved projection up to this accuracy.
    """

    def __init__(self, ctx: Context, rho1: float, rho2: float, dt: float, gx: float = 0.0, gy: float = -0.98,
                 reconstruction: str = "PLIC", tol: float = 1e-13, max_iter: int = 200000,
                 precond: int = L.PRECOND_JACOBI):
        self.ctx = ctx
        self.rho1, self.rho2, self.dt = float(rho1), float(rho2), float(dt)
        self.gx, self.gy = float(gx), float(gy)
        self.recon = reconstruction
        self.tol, self.max_iter, self.precond = float(tol), int(max_iter), int(precond)
        self.C, self.u, self.v, self.x = [], [], [], []
        self.iters = []
        ctx.reserve(L.RESERVE_ADJOINT | (L.RESERVE_MULTIGRID if self.precond in (L.PRECOND_MG, L.PRECOND_SGS_RB)
                                         else 0))

    def _props(self, C):
        ctx = self.ctx
        Cw = C.clone()
        return ctx.update_physical_properties(Cw, self.rho1, self.rho2, 1.0, 1.0)

    def _predictor(self, n):
        first = n == 0
        u_nm1 = None if first else self.u[n - 1]
        v_nm1 = None if first else self.v[n - 1]
        return self.ctx.inertial_predictor(first, self.dt, self.gx, self.gy, self.u[n], self.v[n], u_nm1, v_nm1)

    def forward(self, C0, u0, v0, nsteps: int) -> float:
        """Runs ``nsteps`` steps from (C0, u0, v0) (halo fields; C0's halo is
        refilled from its interior) and returns J = KE(C^N, u^N, v^N)."""
        ctx, dt = self.ctx, self.dt
        C = ctx.fill_vof_halo(C0.clone())
        self.C, self.u, self.v, self.x, self.iters = [C.clone()], [u0.clone()], [v0.clone()], [None], []
        u, v = u0.clone(), v0.clone()
        p = ctx.new_field()
        for n in range(int(nsteps)):
            props = ctx.advect_vof(C, u, v, dt, reconstruction_method=self.recon, two_phase=True, rho1=self.rho1,
                                   rho2=self.rho2, mu1=1.0, mu2=1.0, step_count=n)
            us, vs = self._predictor(n)
            b = ctx.poisson_assemble(us, vs, props["rho_u"], props["rho_v"], dt)
            x = ctx.new_work()
            res = ctx.poisson_solve(b, x, L.LINEAR_CG, self.precond, self.tol, self.max_iter)
            if res.status != L.OK:
                raise L.CovofError(res.status, f"KineticEnergyAdjoint.forward step {n}",
                                   f"pressure solve: {res.iters} iterations, relres {res.relres:.3e}")
            self.iters.append(res.iters)
            ctx.sp_pressure(x, p)
            ctx.correct_velocity(p, props["rho_u"], props["rho_v"], dt, us, vs)
            u, v = us, vs
            self.C.append(C.clone())
            self.u.append(u.clone())
            self.v.append(v.clone())
            self.x.append(x)
        return ctx.kinetic_energy(C, u, v, self.rho1, self.rho2)

    def backward(self):
        """(dJ/dC0, dJ/du0, dJ/dv0) of the last forward run (halo fields; the
        C0 gradient's halo entries are folded onto the interior cells the
        halo copies, so they are zero)."""
        ctx, dt = self.ctx, self.dt
        N = len(self.C) - 1
        if N < 1:
            raise ValueError("backward needs a forward run of at least one step")
        _, Cbar, Wu, Wv = ctx.kinetic_energy(self.C[N], self.u[N], self.v[N], self.rho1, self.rho2, grad=True)
        Au, Av = ctx.new_field(), ctx.new_field()        # dJ/du^n, dJ/dv^n being accumulated
        Pu, Pv = ctx.new_field(), ctx.new_field()        # dJ/du^{n-1}, dJ/dv^{n-1} (predictor's n-1 term)
        for n in range(N - 1, -1, -1):
            props = self._props(self.C[n + 1])
            us, vs = self._predictor(n)
            ctx.poisson_assemble(us, vs, props["rho_u"], props["rho_v"], dt)
            dus, dvs, dlr, res = ctx.projection_adjoint(props["rho"], props["rho_u"], props["rho_v"], dt,
                                                        self.x[n + 1], Wu, Wv, precond=self.precond, tol=self.tol,
                                                        max_iter=self.max_iter)
            if res.status != L.OK:
                raise L.CovofError(res.status, f"KineticEnergyAdjoint.backward step {n}",
                                   f"transpose solve: {res.iters} iterations, relres {res.relres:.3e}")
            ctx.properties_adjoint(self.C[n + 1], self.rho1, self.rho2, dlr, Cbar)
            ctx.inertial_predictor_adjoint(n == 0, dus, dvs, Au, Av, None if n == 0 else Pu,
                                           None if n == 0 else Pv)
            Cbar, _, _ = ctx.vof_advect_adjoint(self.C[n], self.u[n], self.v[n], dt, Cbar, n, self.recon,
                                                two_phase=True, fold_input_halo=(n == 0), dJ_du=Au, dJ_dv=Av)
            Wu, Wv = Au, Av
            Au, Av = Pu, Pv
            Pu, Pv = ctx.new_field(), ctx.new_field()
        return Cbar, Wu, Wv


class _KineticEnergy(torch.autograd.Function):
    @staticmethod
    def forward(fctx, C0, u0, v0, runner: KineticEnergyAdjoint, nsteps: int):
        J = runner.forward(C0, u0, v0, nsteps)
        fctx.runner = runner
        return torch.tensor(J, dtype=torch.float64, device=C0.device)

    @staticmethod
    def backward(fctx, gJ):
        dC, du, dv = fctx.runner.backward()
        return gJ * dC, gJ * du, gJ * dv, None, None


def kinetic_energy_after(runner: KineticEnergyAdjoint, C0, u0, v0, nsteps: int) -> torch.Tensor:
    """J = KE after ``nsteps`` steps from (C0, u0, v0) as a torch scalar whose
    backward is the checkpointed adjoint (gradients for C0, u0 and v0)."""
    return _KineticEnergy.apply(C0, u0, v0, runner, int(nsteps))
