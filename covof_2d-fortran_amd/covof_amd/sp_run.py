"""The time step of a two-phase run on the GPU: standard projection or the
monolithic (fully coupled) system.

Host-side mirror of the body of the reference's time loop for
``SOLVER_STANDARD_PROJECTION`` with two-phase VOF (src/module_simulation.f90:
495-559 and ``sim_solve_standard_projection``, :2424-2666), chaining the
library's entry points in the reference's order:

1. the n-1 / old swaps of the velocity (:503-523);
2. ``advect_vof``: Weymouth-Yue transport, ``update_physical_properties`` and
   the curvature + CSF forces from C (src/module_interface_vof.f90:792-825);
3. ``apply_velocity_bc`` (:541; ``apply_vof_bc`` after it changes nothing,
   the transport leaves the VOF halo filled);
4. the momentum block (velocity predictor, :2454-2469);
5. the Poisson block (:2477-2588) and the velocity correction (:2590-2654);
6. ``sync_u`` / ``sync_v`` + ``apply_boundary_conditions`` (:2657-2660).

With ``SOLVER_MONOLITHIC`` steps 4-6 are the one-pass Uzawa loop of sim_run
(:550-583): blocks, right-hand side, pack (u, v, p), one Krylov solve of the
coupled system, unpack into u, v, p.

Supported configuration: the momentum block's (UNIFORM profiles, no
obstacles, no energy equation, no markers).  SVOF and diffusion smoothing
run in advect_vof's order: SVOF smoothing of C, the properties from C_s,
diffusion smoothing, the CSF normals from C_tilde (else C_s, else C)
(src/module_interface_vof.f90:810-824, :681-689).
Every field is a float64 CUDA tensor in the reference's halo layout; all the
arithmetic runs in libcovof_hip.
"""
from __future__ import annotations

from . import _lib as L
from .context import Context


class StandardProjectionRun:
    """State and time step of a two-phase standard-projection run on one
    context.  Initial properties are the reference's reset values (rho 1,
    mu 1e-3) followed by one ``update_physical_properties``; pressure starts
    at zero.

    ``poisson`` = (linear solver, preconditioner, tol, max_iter, precision) of
    the Poisson block; the reference hard-codes CG + SSOR, 1e-7, 1000, FP64
    (src/module_simulation.f90:2437-2438).  ``predictor`` = (linear solver,
    preconditioner, tol, max_iter) of the momentum block (config_t's
    linear_solver, precond_type, tol, max_iter); with ``solver_type``
    SOLVER_MONOLITHIC it is the coupled system's and ``poisson`` is unused.
    ``curvature``: "normals" (the reference's) or "height_function" (see
    Context.compute_curvature_and_csf).  ``reduction``: the summation order
    of the FP64 inner products (Context.set_reduction_order); with
    REDUCTION_REFERENCE every step is the reference's arithmetic bit for bit.
    ``svof``: the SVOF filter width (cfg%svof_filter_width) with SVOF
    smoothing on, else None; ``diffusion``: the diffusion-smoothing step
    count with diffusion smoothing on, else None.
    """

    FIELDS = ("C", "u", "v", "p", "rho", "rho_u", "rho_v", "mu", "mu_node", "curvature", "fx_sigma", "fy_sigma",
              "u_old", "v_old", "u_nm1", "v_nm1")

    def __init__(self, ctx: Context, C, u, v, rho1, rho2, mu1, mu2, sigma, time_scheme=L.TIME_SCHEME_GEAR,
                 convective_scheme=L.CONV_UPWIND, convective_linearization=L.CONV_LIN_PICARD, gx=0.0, gy=0.0,
                 bc_u=(0.0,) * 4, bc_v=(0.0,) * 4, reconstruction_method="PLIC",
                 predictor=(L.LINEAR_CG, L.PRECOND_JACOBI, 1e-10, 500),
                 poisson=(L.LINEAR_CG, L.PRECOND_SSOR, 1e-7, 1000, L.PRECISION_FP64), convection=True,
                 solver_type=L.SOLVER_STANDARD_PROJECTION, curvature="normals", reduction=None, svof=None,
                 diffusion=None):
        self.ctx = ctx
        if reduction is not None:
            ctx.set_reduction_order(reduction)
        self.rho1, self.rho2, self.mu1, self.mu2, self.sigma = (float(a) for a in (rho1, rho2, mu1, mu2, sigma))
        self.opts = dict(time_scheme=time_scheme, enable_convection=bool(convection), convective_scheme=convective_scheme,
                         convective_linearization=convective_linearization, gx=gx, gy=gy, bc_u=tuple(bc_u),
                         bc_v=tuple(bc_v), solver_type=int(solver_type))
        self.mono = int(solver_type) == L.SOLVER_MONOLITHIC
        self.curvature = curvature
        self.recon = reconstruction_method
        self.predictor = predictor
        self.poisson = poisson
        self.svof = None if svof is None else float(svof)
        self.diffusion = None if diffusion is None else int(diffusion)
        # the work arrays of every block the step uses, made now: the steps
        # themselves allocate nothing on the device
        parts = L.RESERVE_MOMENTUM | L.RESERVE_SURFACE
        pcs = {int(predictor[1]), int(poisson[1])}
        if pcs & {L.PRECOND_MG, L.PRECOND_SGS_RB}:
            parts |= L.RESERVE_MULTIGRID
        if L.PRECOND_SSOR in pcs:
            parts |= L.RESERVE_SGS
        ctx.reserve(parts)
        f = {}
        f["C"] = ctx._field(C, "C").clone()
        f["u"] = ctx._field(u, "u").clone()
        f["v"] = ctx._field(v, "v").clone()
        for k in ("p", "curvature", "fx_sigma", "fy_sigma", "u_old", "v_old", "u_nm1", "v_nm1"):
            f[k] = ctx.new_field(0.0)
        for k in ("rho", "rho_u", "rho_v"):
            f[k] = ctx.new_field(1.0)
        for k in ("mu", "mu_node"):
            f[k] = ctx.new_field(1.0e-3)
        if self.svof is not None:
            f["C_s"] = ctx.new_field(1.0)
        if self.diffusion is not None:
            f["C_tilde"] = ctx.new_field(0.0)
        self.f = f
        self.props = {k: f[k] for k in ("rho", "rho_u", "rho_v", "mu", "mu_node")}
        # the end of set_initial_conditions (src/module_interface_vof.f90:368-374)
        if self.svof is not None:
            ctx.smooth_svof_helmholtz(f["C"], self.svof, C_s=f["C_s"])
        if self.diffusion is not None:
            ctx.smooth_diffusion(f["C"], self.diffusion, C_tilde=f["C_tilde"])
        ctx.update_physical_properties(f["C"], self.rho1, self.rho2, self.mu1, self.mu2, props=self.props,
                                       C_s=f.get("C_s"))
        self.nstep = 0
        ctx.step_count = 0

    def cfg(self, has_nm1: bool) -> L.MomentumCfg:
        return Context.momentum_cfg(has_nm1=has_nm1, **self.opts)

    def step(self, dt: float):
        """One time step; returns (predictor iterations, Poisson iterations,
        predictor status, Poisson status) -- monolithic: (coupled solve
        iterations, 0, its status, OK)."""
        ctx, f = self.ctx, self.f
        self.nstep += 1
        has_nm1 = self.nstep > 1
        if has_nm1:
            f["u_nm1"].copy_(f["u_old"])
            f["v_nm1"].copy_(f["v_old"])
        else:
            f["u_nm1"].copy_(f["u"])
            f["v_nm1"].copy_(f["v"])
        f["u_old"].copy_(f["u"])
        f["v_old"].copy_(f["v"])
        # advect_vof: transport, [SVOF smoothing], properties, [diffusion
        # smoothing], then curvature and CSF
        if self.svof is None:
            ctx.advect_vof(f["C"], f["u"], f["v"], dt, reconstruction_method=self.recon, two_phase=True,
                           rho1=self.rho1, rho2=self.rho2, mu1=self.mu1, mu2=self.mu2, props=self.props)
        else:
            ctx.advect_vof(f["C"], f["u"], f["v"], dt, reconstruction_method=self.recon)
            ctx.smooth_svof_helmholtz(f["C"], self.svof, C_s=f["C_s"])
            ctx.update_physical_properties(f["C"], self.rho1, self.rho2, self.mu1, self.mu2, props=self.props,
                                           C_s=f["C_s"])
        G = f.get("C_s")
        if self.diffusion is not None:
            ctx.smooth_diffusion(f["C"], self.diffusion, C_tilde=f["C_tilde"])
            G = f["C_tilde"]
        ctx.compute_curvature_and_csf(f["C"], self.sigma, G=G, enable=self.sigma > 0.0,
                                      out={k: f[k] for k in ("curvature", "fx_sigma", "fy_sigma")},
                                      method=self.curvature)
        cfg = self.cfg(has_nm1)
        ctx.apply_velocity_bc(cfg, f["u"], f["v"])
        mom = {k: f[k] for k in ("u", "v", "u_old", "v_old", "u_nm1", "v_nm1", "rho_u", "rho_v", "mu", "mu_node",
                                 "fx_sigma", "fy_sigma")}
        if self.mono:
            # the coupled solve of u, v, p (unpacked with the velocity BC)
            rm = ctx.solve_monolithic_step(cfg, mom, f["p"], dt, *self.predictor)
            return rm.iters, 0, rm.status, L.OK
        # the momentum block
        rp = ctx.solve_standard_projection_predictor(cfg, mom, dt, *self.predictor)
        # the Poisson block, the correction, sync + BC
        solver, precond, tol, maxit, precision = self.poisson
        rq = ctx.solve_standard_projection_poisson(f["u"], f["v"], f["rho_u"], f["rho_v"], f["p"], dt, solver,
                                                   precond, tol, maxit, precision)
        self.last_poisson = rq  # iterations, relres (the true FP64 residual for MIXED), status
        ctx.correct_velocity(f["p"], f["rho_u"], f["rho_v"], dt, f["u"], f["v"])
        ctx.apply_velocity_bc(cfg, f["u"], f["v"])
        return rp.iters, rq.iters, rp.status, rq.status

    def kinetic_energy(self) -> float:
        """0.5 sum(rho |u|^2) dx dy at cell centres (diagnostic)."""
        ctx, f = self.ctx, self.f
        nx, ny = ctx.nx, ctx.ny
        uc = 0.5 * (f["u"][1:ny + 1, 1:nx + 1] + f["u"][1:ny + 1, 2:nx + 2])
        vc = 0.5 * (f["v"][1:ny + 1, 1:nx + 1] + f["v"][2:ny + 2, 1:nx + 1])
        dxdy = (ctx.lx / nx) * (ctx.ly / ny)
        return float((0.5 * f["rho"][1:ny + 1, 1:nx + 1] * (uc * uc + vc * vc)).sum().item() * dxdy)

