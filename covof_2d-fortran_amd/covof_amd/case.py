"""A reference JSON case run on the GPU: the drop-in for ``covof2d_fortran case.json``.

``load_config`` reads the keys of ``load_config_from_json``
(src/module_config.f90:255-727) with the reference's defaults and string
selectors; ``run_case`` reproduces ``sim_init`` (src/module_simulation.f90:
96-351: mesh, reset fields, ``set_initial_conditions`` of
src/module_interface_vof.f90:47-380, velocity BC, writer) and ``sim_run``
(:416-755: the time loop with its step-0 output, the end-time clamp, the
history row every step, the VTK files every ``vtk_interval`` steps and at the
end) for a two-phase standard-projection case, with every time step on the
GPU (``StandardProjectionRun``) and the reference's output files
(``covof_amd.output``).

Supported: mode NAVIER_STOKES, solver STANDARD_PROJECTION, two-phase VOF with
WEYMOUTH_YUE + PLIC / ELVIRA (with SVOF / diffusion smoothing), no markers, no
energy equation, incompressible, no obstacles, fixed dt, UNIFORM boundary
profiles, the predictor's CG / BiCGSTAB(2) with NONE / JACOBI, and the
initial conditions UNIFORM, DROPLET, CYLINDER and BUBBLE.  Anything else
raises ``ValueError`` naming the key.  The Poisson block is the reference's
(CG + SGS, 1e-7, 1000 iterations) unless ``poisson`` says otherwise.  With
``--reference-order`` the FP64 inner products are summed in the reference's
order and the run reproduces the reference program's output files exactly.

    python -m covof_amd.case case.json [--poisson-mg-mixed] [--reference-order]
"""
from __future__ import annotations

import json
import math
import os
import sys
import time as _time

import numpy as np

from . import _lib as L

_BC = {"DIRICHLET": L.BC_DIRICHLET, "NEUMANN": L.BC_NEUMANN, "PERIODIC": L.BC_PERIODIC,
       "FREE_SLIP": L.BC_FREESLIP}


def _get(d: dict, path: str, default):
    cur = d
    for part in path.split("."):
        if not isinstance(cur, dict) or part not in cur:
            return default
        cur = cur[part]
    return cur


def load_config(path: str) -> dict:
    """The config_t fields the GPU run reads, from a reference JSON case
    (key names, fallbacks and defaults of load_config_from_json)."""
    with open(path) as f:
        j = json.load(f)
    g = lambda p, d: _get(j, p, d)  # noqa: E731
    c = {}
    c["case_name"] = g("output.case_name", g("case_name", "covof_case"))
    c["mode"] = g("mode", "NAVIER_STOKES")
    c["nx"], c["ny"] = int(g("mesh.nx", 64)), int(g("mesh.ny", 64))
    c["lx"], c["ly"] = float(g("mesh.lx", 1.0)), float(g("mesh.ly", 1.0))
    c["enable_two_phase"] = bool(g("physics.enable_two_phase", False))
    c["enable_surface_tension"] = bool(g("physics.enable_surface_tension", False))
    c["enable_convection"] = bool(g("physics.enable_convection", True))
    c["sigma"] = float(g("physics.sigma", 0.0))
    c["gx"], c["gy"] = float(g("physics.gx", 0.0)), float(g("physics.gy", 0.0))
    cs = g("physics.convective_scheme", "UPWIND")
    c["convective_scheme"] = {"CENTERED": L.CONV_CENTERED, "TVD_SUPERBEE": L.CONV_TVD_SUPERBEE}.get(cs, L.CONV_UPWIND)
    cl = g("physics.convective_linearization", "PICARD")
    if cl == "PICARD":
        cl = g("convective_linearization", "PICARD")
    c["convective_linearization"] = {"ADAMS_BASHFORTH": L.CONV_LIN_ADAMS_BASHFORTH, "AB2": L.CONV_LIN_ADAMS_BASHFORTH,
                                     "NEWTON": L.CONV_LIN_NEWTON}.get(cl, L.CONV_LIN_PICARD)
    c["enable_svof_smoothing"] = bool(g("physics.enable_svof_smoothing", g("enable_svof_smoothing", False)))
    c["enable_diffusion_smoothing"] = bool(g("physics.enable_diffusion_smoothing", False))
    # svof_filter_width: physics.svof_filter_width, else the top-level key when that left the default
    # (src/module_config.f90:340-343); diffusion_smoothing_steps (:381)
    fw = float(g("physics.svof_filter_width", 1.0))
    c["svof_filter_width"] = float(g("svof_filter_width", 1.0)) if fw == 1.0 else fw
    c["diffusion_smoothing_steps"] = int(g("physics.diffusion_smoothing_steps", 3))
    c["viscosity_model"] = g("physics.viscosity_model", "CONSTANT")
    c["rho1"] = float(g("physics.fluid1.rho", 1.0))
    c["mu1"] = float(g("physics.fluid1.mu", 1.0e-3))
    c["rho2"] = float(g("physics.fluid2.rho", c["rho1"]))
    c["mu2"] = float(g("physics.fluid2.mu", c["mu1"]))
    c["compressible"] = bool(g("physics.compressible", False)) or g("physics.flow_regime", "INCOMPRESSIBLE") != \
        "INCOMPRESSIBLE"
    c["enable_ssp_rk3_momentum"] = bool(g("physics.enable_ssp_rk3_momentum", c["compressible"]))
    c["enable_energy"] = bool(g("thermal.enable_energy", g("physics.enable_energy", False)))
    ts = g("time.time_scheme", "EULER")
    if ts == "EULER":
        ts = g("time_scheme", "EULER")
    c["time_scheme"] = L.TIME_SCHEME_GEAR if ts in ("GEAR", "BDF2") else L.TIME_SCHEME_EULER
    c["dt"] = float(g("time.dt", 0.01))
    c["t_final"] = float(g("time.t_final", 1.0))
    c["adaptive_dt"] = bool(g("time.adaptive_dt", False))
    for side in ("left", "right", "bottom", "top"):
        p = f"boundary_conditions.{side}"
        c[f"bc_{side}"] = _BC.get(g(p + ".type", "DIRICHLET"), L.BC_DIRICHLET)
        u = float(g(p + ".u", 0.0))
        u = float(g(p + ".u_mean", u))
        c[f"bc_{side}_u"] = float(g(p + ".u_val", u))
        c[f"bc_{side}_v"] = float(g(p + ".v_val", float(g(p + ".v", 0.0))))
        c[f"bc_{side}_profile"] = g(p + ".profile", "UNIFORM")
    c["obstacle"] = g("obstacle", None) is not None or g("obstacles", None) is not None
    ic = g("initial_condition.type", "") or g("initial_conditions.type", "UNIFORM")
    c["ic_type"] = "UNIFORM" if ic == "SINGLE_PHASE" else ic

    def ic_real(k, d):
        v = float(g("initial_condition." + k, d))
        return float(g("initial_conditions." + k, v))

    c["ic_u"], c["ic_v"] = ic_real("u", 0.0), ic_real("v", 0.0)
    c["ic_center_x"], c["ic_center_y"] = ic_real("center_x", 0.5), ic_real("center_y", 0.5)
    c["ic_radius"] = ic_real("radius", 0.25)
    c["ic_axis_a"], c["ic_axis_b"] = ic_real("axis_a", 0.0), ic_real("axis_b", 0.0)
    st = g("solver.type", "AUGMENTED_LAGRANGIAN")
    c["solver_type"] = {"PROJECTION": "STANDARD_PROJECTION", "FULLY_COUPLED": "MONOLITHIC",
                        "FC": "MONOLITHIC"}.get(st, st)
    ls = g("solver.linear_solver", "BICGSTAB2")
    c["linear_solver"] = {"CG": L.LINEAR_CG, "CONJUGATE_GRADIENT": L.LINEAR_CG}.get(
        ls, L.LINEAR_BICGSTAB2 if ls == "BICGSTAB2" else ls)
    c["tol"] = float(g("solver.tol", 1.0e-5))
    c["max_iter"] = int(g("solver.max_iter", 300))
    use_precond = bool(g("solver.use_precond", True))
    pc = g("solver.precond_type", "") or g("solver.preconditioner", "JACOBI")
    pcode = {"JACOBI": L.PRECOND_JACOBI, "DIAGONAL": L.PRECOND_JACOBI, "NONE": L.PRECOND_NONE,
             "SSOR": L.PRECOND_SSOR, "GAUSS_SEIDEL": L.PRECOND_SSOR}.get(pc, pc)
    if pc in ("BLOCK_TRIANGULAR", "SCHUR", "BT_SBP", "BT-SBP", "BLOCK", "BLOCK_PRECONDITIONER"):
        pcode = L.PRECOND_BLOCK_TRIANGULAR
    c["precond"] = pcode if use_precond else L.PRECOND_NONE
    c["advection_method"] = g("vof.advection_method", "WEYMOUTH_YUE")
    rm = g("vof.reconstruction_method", "PLIC")
    c["reconstruction_method"] = "PLIC" if rm in ("NONE", "") else rm
    c["use_markers"] = bool(g("markers.enabled", g("vof.use_markers", False)))
    c["output_dir"] = g("output.directory", g("output.output_dir", "output"))
    c["vtk_interval"] = int(g("output.vtk_interval", 50))
    c["log_interval"] = int(g("output.log_interval", 10))
    c["write_csv"] = bool(g("output.write_csv_history", True))
    return c


def check_supported(c: dict) -> None:
    """Raises ValueError naming the first setting the GPU run does not cover."""
    rules = [
        (c["mode"] == "NAVIER_STOKES", "mode"), (c["solver_type"] in ("STANDARD_PROJECTION", "MONOLITHIC"), "solver.type"),
        (c["enable_two_phase"], "physics.enable_two_phase"),
        (c["advection_method"] == "WEYMOUTH_YUE", "vof.advection_method"),
        (c["reconstruction_method"] in ("PLIC", "ELVIRA"), "vof.reconstruction_method"),
        (not c["use_markers"], "markers"),
        (c["viscosity_model"] == "CONSTANT", "physics.viscosity_model"),
        (not c["compressible"], "physics.compressible"), (not c["enable_ssp_rk3_momentum"], "ssp_rk3"),
        (not c["enable_energy"], "enable_energy"), (not c["adaptive_dt"], "time.adaptive_dt"),
        (not c["obstacle"], "obstacle"),
        (c["linear_solver"] in (L.LINEAR_CG, L.LINEAR_BICGSTAB2), "solver.linear_solver"),
        (c["precond"] in (L.PRECOND_NONE, L.PRECOND_JACOBI, L.PRECOND_SSOR, L.PRECOND_BLOCK_TRIANGULAR),
         "solver.preconditioner"),
        (c["ic_type"] in ("UNIFORM", "DROPLET", "CYLINDER", "BUBBLE"), "initial_condition.type"),
        (c["bc_top"] != L.BC_NEUMANN, "boundary_conditions.top (NEUMANN)"),
    ]
    for side in ("left", "right", "bottom", "top"):
        rules.append((c[f"bc_{side}_profile"] == "UNIFORM", f"boundary_conditions.{side}.profile"))
    for ok, key in rules:
        if not ok:
            raise ValueError(f"not supported on the GPU path: {key}")


def initial_volume_fraction(c: dict) -> np.ndarray:
    """set_initial_conditions' C over the halo grid (src/module_interface_vof.f90:
    60-260): UNIFORM, and DROPLET / CYLINDER / BUBBLE by 16 x 16 sub-cell
    counts or the smoothed-distance profile, as the reference chooses."""
    nx, ny, dx, dy = c["nx"], c["ny"], c["lx"] / c["nx"], c["ly"] / c["ny"]
    C = np.ones((ny + 2, nx + 2))
    t = c["ic_type"]
    if t == "UNIFORM":
        return C
    xc, yc, r, a, b = c["ic_center_x"], c["ic_center_y"], c["ic_radius"], c["ic_axis_a"], c["ic_axis_b"]
    ellipse = a > 0.0 and b > 0.0
    eps = 1.5 * max(dx, dy)
    sub = (np.arange(16) + 0.5)
    gas = t in ("CYLINDER", "BUBBLE")
    sampled = (c["use_markers"] or ellipse) if t == "DROPLET" else \
        (c["use_markers"] or t == "CYLINDER" or c["reconstruction_method"] == "PLIC" or ellipse)
    for j in range(ny + 2):
        for i in range(nx + 2):
            if sampled:
                xs = float(i - 1) * dx + sub * (dx / 16.0)   # x_face(i) + (si + 0.5) dx / 16
                ys = float(j - 1) * dy + sub * (dy / 16.0)
                X, Y = np.meshgrid(xs, ys)
                if ellipse:
                    q = ((X - xc) / a) ** 2 + ((Y - yc) / b) ** 2
                    hit = (q > 1.0) if gas else (q <= 1.0)
                else:
                    d = np.hypot(X - xc, Y - yc)
                    hit = (d > r) if gas else (d <= r)
                C[j, i] = float(np.count_nonzero(hit)) / 256.0
            else:
                x, y = (float(i - 1) + 0.5) * dx, (float(j - 1) + 0.5) * dy
                dist = math.sqrt((x - xc) ** 2 + (y - yc) ** 2) - r
                if dist < -eps:
                    C[j, i] = 0.0 if gas else 1.0
                elif dist > eps:
                    C[j, i] = 1.0 if gas else 0.0
                elif gas:
                    C[j, i] = 0.5 + dist / (2.0 * eps) + (1.0 / (2.0 * math.pi)) * math.sin(math.pi * dist / eps)
                else:
                    C[j, i] = 0.5 - dist / (2.0 * eps) - (1.0 / (2.0 * math.pi)) * math.sin(math.pi * dist / eps)
    return C


def _vof_halo(C: np.ndarray, c: dict) -> None:
    """exchange_halos_scalar + apply_vof_bc on one rank (no JET_INLET)."""
    nx, ny = c["nx"], c["ny"]
    if L.BC_PERIODIC in (c["bc_left"], c["bc_right"]):
        C[:, 0] = C[:, nx]
        C[:, nx + 1] = C[:, 1]
    if L.BC_PERIODIC in (c["bc_bottom"], c["bc_top"]):
        C[0, :] = C[ny, :]
        C[ny + 1, :] = C[1, :]
    if c["bc_left"] == L.BC_PERIODIC:
        C[1:ny + 1, 0] = C[1:ny + 1, nx]
        C[1:ny + 1, nx + 1] = C[1:ny + 1, 1]
    else:
        C[1:ny + 1, 0] = C[1:ny + 1, 1]
        C[1:ny + 1, nx + 1] = C[1:ny + 1, nx]
    if c["bc_bottom"] == L.BC_PERIODIC:
        C[0, 1:nx + 1] = C[ny, 1:nx + 1]
        C[ny + 1, 1:nx + 1] = C[1, 1:nx + 1]
    else:
        C[0, 1:nx + 1] = C[1, 1:nx + 1]
        C[ny + 1, 1:nx + 1] = C[ny, 1:nx + 1]


def run_case(c: dict, poisson=(L.LINEAR_CG, L.PRECOND_SSOR, 1e-7, 1000, L.PRECISION_FP64), log=print,
             reduction=None):
    """sim_init + sim_run of a supported case on the GPU; returns the run.
    ``reduction``: the summation order of the inner products
    (Context.set_reduction_order; None: the default tree sums)."""
    import torch

    from .context import Context
    from .output import HistoryWriter, write_step
    from .sp_run import StandardProjectionRun
    check_supported(c)
    nx, ny = c["nx"], c["ny"]
    bc = (c["bc_left"], c["bc_right"], c["bc_bottom"], c["bc_top"])
    ctx = Context(nx, ny, c["lx"], c["ly"], bc=bc)
    C = initial_volume_fraction(c)
    _vof_halo(C, c)
    dev = ctx.device
    u = torch.full((ny + 2, nx + 2), c["ic_u"], dtype=torch.float64, device=dev)
    v = torch.full((ny + 2, nx + 2), c["ic_v"], dtype=torch.float64, device=dev)
    run = StandardProjectionRun(
        ctx, torch.tensor(C, device=dev), u, v, c["rho1"], c["rho2"], c["mu1"], c["mu2"],
        c["sigma"] if c["enable_surface_tension"] else 0.0, time_scheme=c["time_scheme"],
        convective_scheme=c["convective_scheme"], convective_linearization=c["convective_linearization"],
        gx=c["gx"], gy=c["gy"], bc_u=tuple(c[f"bc_{s}_u"] for s in ("left", "right", "bottom", "top")),
        bc_v=tuple(c[f"bc_{s}_v"] for s in ("left", "right", "bottom", "top")),
        reconstruction_method=c["reconstruction_method"],
        predictor=(c["linear_solver"], c["precond"], c["tol"], c["max_iter"]), poisson=poisson,
        convection=c["enable_convection"],
        svof=c["svof_filter_width"] if c["enable_svof_smoothing"] else None,
        diffusion=c["diffusion_smoothing_steps"] if c["enable_diffusion_smoothing"] else None, reduction=reduction,
        solver_type=L.SOLVER_MONOLITHIC if c["solver_type"] == "MONOLITHIC" else L.SOLVER_STANDARD_PROJECTION)
    f = run.f
    # the end of set_initial_conditions: CSF of the initial C; then sim_init's velocity BC
    G = f.get("C_tilde", f.get("C_s"))  # the normals' field of compute_curvature_and_csf (:681-689)
    ctx.compute_curvature_and_csf(f["C"], run.sigma, G=G, enable=run.sigma > 0.0,
                                  out={k: f[k] for k in ("curvature", "fx_sigma", "fy_sigma")})
    ctx.apply_velocity_bc(run.cfg(False), f["u"], f["v"])
    out_dir, case = c["output_dir"], c["case_name"]
    hist = HistoryWriter(ctx, out_dir, case)
    fields = lambda: {k: f[k] for k in ("C", "C_s", "p", "rho", "mu", "curvature", "u", "v") if k in f}  # noqa: E731
    write_step(ctx, out_dir, case, 0, fields())
    hist.write(0, 0.0, c["dt"], f["C"], f["u"], f["v"], f["p"], f["rho"])
    t, step, tf = 0.0, 0, c["t_final"]
    t0 = _time.perf_counter()
    while t < tf - 1.0e-12 * (abs(tf) + 1.0):
        step += 1
        dt = c["dt"]
        if t + dt > tf:
            dt = tf - t
        if dt < 1.0e-12:
            break
        iu, ip, _, _ = run.step(dt)
        t = t + dt
        if c["write_csv"]:
            vals = hist.write(step, t, dt, f["C"], f["u"], f["v"], f["p"], f["rho"])
        else:
            vals = ctx.history_row(f["C"], f["u"], f["v"], f["p"], f["rho"])
        if c["log_interval"] > 0 and step % c["log_interval"] == 0:
            log(f"{step:8d}{t:14.5f}{dt:12.3E}{iu + ip:14d}{vals[4]:16.4E}{vals[5]:16.4E}")
        if step % c["vtk_interval"] == 0 or t >= tf - 1.0e-10 * (abs(tf) + 1.0):
            write_step(ctx, out_dir, case, step, fields())
    torch.cuda.synchronize()
    run.wall_s = _time.perf_counter() - t0
    run.steps = step
    return run


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print(__doc__)
        return 2
    c = load_config(argv[0])
    poisson = (L.LINEAR_CG, L.PRECOND_MG, 1e-10, 20000, L.PRECISION_MIXED) if "--poisson-mg-mixed" in argv else \
        (L.LINEAR_CG, L.PRECOND_SSOR, 1e-7, 1000, L.PRECISION_FP64)
    run = run_case(c, poisson=poisson, reduction=L.REDUCTION_REFERENCE if "--reference-order" in argv else None)
    print(f"{run.steps} steps in {run.wall_s:.2f} s; output in {os.path.abspath(c['output_dir'])}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
