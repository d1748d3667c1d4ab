"""The CPU oracle against the golden vectors of the compiled reference.

Fixtures (tests/golden/*.npz) hold inputs and outputs of the reference's own
public procedures (cut_box_area, compute_alpha, reconstruct_interface,
advect_vof + update_physical_properties), built with -ffp-contract=off.  The
oracle must reproduce them bit-for-bit.
"""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_geom_bitwise(orc):
    g = _load("geom.npz")
    n = len(g["xmin"])
    area = np.array([orc.cut_box_area(g["xmin"][k], g["xmax"][k], g["ymin"][k], g["ymax"][k],
                                      g["nx"][k], g["ny"][k], g["alpha"][k]) for k in range(n)])
    alpha = np.array([orc.compute_alpha(g["xmin"][k], g["xmax"][k], g["ymin"][k], g["ymax"][k],
                                        g["nx"][k], g["ny"][k], g["c"][k]) for k in range(n)])
    assert np.array_equal(area, g["area"])
    assert np.array_equal(alpha, g["alpha_out"])


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "recon_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_reconstruct_interface_bitwise(orc, path):
    g = np.load(path, allow_pickle=False)
    m = orc.mesh(int(g["nx"]), int(g["ny"]), float(g["lx"]), float(g["ly"]), tuple(g["bc"]))
    sx, sy, sa = orc.reconstruct_interface(m, int(g["method"]), g["C"])
    assert np.array_equal(sx, g["snx"])
    assert np.array_equal(sy, g["sny"])
    assert np.array_equal(sa, g["sal"])


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "advect_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_advect_vof_bitwise(orc, path):
    g = np.load(path, allow_pickle=False)
    nx, ny = int(g["nx"]), int(g["ny"])
    m = orc.mesh(nx, ny, float(g["lx"]), float(g["ly"]), tuple(g["bc"]))
    C = g["C0"].copy()
    sc = 0
    for s in range(int(g["nsteps"])):
        C, sc = orc.wy_step(m, int(g["method"]), C, g["u"], g["v"], float(g["dt"]), sc)
        if int(g["two_phase"]):
            props = orc.update_properties(m, C, float(g["rho1"]), float(g["rho2"]),
                                          float(g["mu1"]), float(g["mu2"]))
        assert np.array_equal(C, g["C_steps"][s]), f"step {s + 1}: max diff {np.abs(C - g['C_steps'][s]).max()}"
    if int(g["two_phase"]):
        for k, name in enumerate(("rho", "rho_u", "rho_v", "mu", "mu_node")):
            assert np.array_equal(props[k], g[name]), name


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "surface_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_surface_steps_bitwise(orc, path):
    """advect_vof with SVOF / diffusion smoothing, curvature and CSF forces:
    C, C_s, C_tilde, curvature, fx_sigma, fy_sigma after every step."""
    g = np.load(path, allow_pickle=False)
    nx, ny = int(g["nx"]), int(g["ny"])
    m = orc.mesh(nx, ny, float(g["lx"]), float(g["ly"]), tuple(g["bc"]))
    C = g["C0"].copy()
    shp = C.shape
    Cs, Ct = np.ones(shp), np.ones(shp)  # the reference's initial values
    fx, fy = np.zeros(shp), np.zeros(shp)
    sc = 0
    for s in range(int(g["nsteps"])):
        C, sc = orc.wy_step(m, int(g["method"]), C, g["u"], g["v"], float(g["dt"]), sc)
        if int(g["svof"]):
            Cs = orc.svof_smooth(m, C, float(g["width"]))
        if int(g["diff"]):
            Ct = orc.smooth_diffusion(m, C, int(g["dsteps"]))
        G = Ct if int(g["diff"]) else (Cs if int(g["svof"]) else C)
        curv, fx, fy = orc.curvature_csf(m, G, C, float(g["sigma"]), int(g["tension"]), fx, fy)
        for name, val in (("C", C), ("C_s", Cs), ("C_tilde", Ct), ("curvature", curv), ("fx_sigma", fx),
                          ("fy_sigma", fy)):
            ref = g[name][s]
            assert np.array_equal(val, ref), f"step {s + 1} {name}: max diff {np.abs(val - ref).max()}"


def _momentum(orc, g):
    nx, ny = int(g["nx"]), int(g["ny"])
    m = orc.mesh(nx, ny, float(g["lx"]), float(g["ly"]), tuple(int(b) for b in g["bc"]))
    opts = dict(time_scheme=int(g["tscheme"]), has_nm1=int(g["has_nm1"]), enable_convection=int(g["conv"]),
                conv_scheme=int(g["scheme"]), conv_lin=int(g["lin"]), gx=float(g["gx"]), gy=float(g["gy"]),
                bc_u=tuple(g["bcu"]), bc_v=tuple(g["bcv"]), solver_type=int(g["stype"]))
    return orc.Momentum(m, {k: g[k] for k in orc.MOM_FIELDS}, opts, float(g["dt"]))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "momentum_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_momentum_block_bitwise(orc, path):
    """assemble_preconditioner_blocks (diag_u, diag_v, and diag_p of the
    monolithic system), assemble_rhs, two apply_system_matrix products in a
    row, apply_preconditioner of the fixture's kind (JACOBI, SSOR or
    BLOCK_TRIANGULAR), and the whole block: the SP predictor (pack,
    solve_linear_system, unpack, sync, BC) or the monolithic solve (u, v, p).
    The oracle's serial inner products are the reference's, so even the solve
    is bit-for-bit."""
    g = np.load(path, allow_pickle=False)
    M = _momentum(orc, g)
    assert np.array_equal(M.diag(), g["diag"])
    assert np.array_equal(M.rhs(), g["rhs"])
    assert np.array_equal(M.matvec(g["x1"]), g["y1"])
    assert np.array_equal(M.matvec(g["x2"]), g["y2"])
    assert np.array_equal(M.precond(int(g["precond"]), g["x1"]), g["z1"])
    args = (int(g["solver"]), int(g["precond"]), float(g["tol"]), int(g["maxit"]))
    if int(g["stype"]) == 2:
        u, v, p, it, _ = M.mono_step(g["p"], *args)
        assert np.array_equal(p, g["p_out"])
    else:
        u, v, it, _ = M.predictor(*args)
    assert it == int(g["iters"])
    assert np.array_equal(u, g["u_out"]) and np.array_equal(v, g["v_out"])


@pytest.mark.parametrize("path", [os.path.join(GOLDEN, "spstep_rising_bubble.npz")],
                         ids=lambda p: os.path.basename(p))
def test_sp_run_bitwise(orc, path):
    """The oracle's restatement of the reference's time-loop body (n-1 / old
    swaps, advect_vof with properties and CSF, velocity BC, momentum block,
    Poisson block, correction, sync + BC) reproduces the compiled
    reference's run bit for bit, iteration counts included."""
    g = np.load(path, allow_pickle=False)
    m = orc.mesh(int(g["nx"]), int(g["ny"]), float(g["lx"]), float(g["ly"]), tuple(int(b) for b in g["bc"]))
    opts = dict(time_scheme=int(g["tscheme"]), conv_scheme=int(g["scheme"]), conv_lin=int(g["lin"]),
                gx=float(g["gx"]), gy=float(g["gy"]), bc_u=tuple(g["bcu"]), bc_v=tuple(g["bcv"]))
    R = orc.SPRun(m, g["C0"], g["u0"], g["v0"], float(g["rho1"]), float(g["rho2"]), float(g["mu1"]),
                  float(g["mu2"]), float(g["sigma"]), opts, int(g["method"]), int(g["solver"]), int(g["precond"]),
                  float(g["tol"]), int(g["maxit"]))
    for s in range(int(g["nsteps"])):
        assert R.step(float(g["dt"])) == int(g["iters"][s])
        if s == 0:
            for a, k in ((R.C, "C1"), (R.u, "u1"), (R.v, "v1"), (R.p, "p1")):
                assert np.array_equal(a, g[k]), k
    for a, k in ((R.C, "C"), (R.u, "u"), (R.v, "v"), (R.p, "p")):
        assert np.array_equal(a, g[k]), k


def test_output_files_text(orc):
    """vtk_writer_t's files: the oracle's history quantities through the
    Fortran ES14.6 edit give the reference's _history.csv character for
    character, and the .vtr / .pvtr builders give the reference's files from
    the same cell data (divergence and cell-centred velocity formed as
    vtk_write_step forms them)."""
    from covof_amd.output import HISTORY_HEADER, fortran_es, pvtr_text, vtr_text
    g = np.load(os.path.join(GOLDEN, "output_writer.npz"), allow_pickle=False)
    nx, ny, lx, ly = int(g["nx"]), int(g["ny"]), float(g["lx"]), float(g["ly"])
    m = orc.mesh(nx, ny, lx, ly, tuple(int(b) for b in g["bc"]))
    lines = [HISTORY_HEADER]
    for k, step in enumerate(g["steps"]):
        f = {n: g[f"{n}_{k}"] for n in ("C", "u", "v", "p", "rho", "mu", "T", "curvature", "C_s")}
        h = orc.history(m, f["C"], f["u"], f["v"], f["p"], f["rho"], f["T"])
        lines.append(f"{int(step)}," + ",".join(fortran_es(x, 14, 6)
                                                for x in (float(g["times"][k]), float(g["dts"][k]), *h)))
        u, v = f["u"], f["v"]
        dx, dy = lx / nx, ly / ny
        arrays = {n: f[n][1:ny + 1, 1:nx + 1] for n in ("C", "C_s", "p", "rho", "mu", "T", "curvature")}
        arrays["div"] = (u[1:ny + 1, 2:nx + 2] - u[1:ny + 1, 1:nx + 1]) / dx + \
            (v[2:ny + 2, 1:nx + 1] - v[1:ny + 1, 1:nx + 1]) / dy
        uc = 0.5 * (u[1:ny + 1, 1:nx + 1] + u[1:ny + 1, 2:nx + 2])
        vc = 0.5 * (v[1:ny + 1, 1:nx + 1] + v[2:ny + 2, 1:nx + 1])
        assert vtr_text(nx, ny, dx, dy, arrays, uc, vc) == str(g[f"vtr_{step}"])
        assert pvtr_text(nx, ny, "case", int(step)) == str(g[f"pvtr_{step}"])
    assert "\n".join(lines) + "\n" == str(g["history"])
    assert fortran_es(1.5e-123, 14, 6) == "  1.500000-123" and fortran_es(-2.0e150, 16, 8) == " -2.00000000+150"


# --------------------------------------------------------------------------
# SVOF properties, the JET_INLET top boundary and solid obstacles in the SP
# Poisson block: the oracle against the compiled reference's fixtures.

@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "props_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_oracle_properties_golden(orc, path):
    g = np.load(path, allow_pickle=False)
    nx, ny = int(g["nx"]), int(g["ny"])
    m = orc.mesh(nx, ny, float(g["lx"]), float(g["ly"]), bc=tuple(int(b) for b in g["bc"]))
    C = g["C0"].copy()
    out = orc.update_properties_cs(m, C, g["C_s"] if int(g["svof"]) else None, float(g["rho1"]), float(g["rho2"]),
                                   float(g["mu1"]), float(g["mu2"]))
    assert np.array_equal(C, g["C"])
    for k, a in zip(("rho", "rho_u", "rho_v", "mu", "mu_node"), out):
        # the reference initialises untouched entries differently (fields%allocate): compare what it writes
        ny2, nx2 = a.shape
        sl = {"rho": np.s_[:, :], "mu": np.s_[:, :], "rho_u": np.s_[1:ny2 - 1, 1:], "rho_v": np.s_[1:, 1:nx2 - 1],
              "mu_node": np.s_[1:, 1:]}[k]
        assert np.array_equal(a[sl], g[k][sl]), k


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "jet_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_oracle_jet_inlet_golden(orc, path):
    g = np.load(path, allow_pickle=False)
    nx, ny = int(g["nx"]), int(g["ny"])
    m = orc.mesh(nx, ny, float(g["lx"]), float(g["ly"]), bc=tuple(int(b) for b in g["bc"]))
    orc.set_top_jet(True, float(g["jet_center"]), float(g["jet_width"]))
    try:
        C, sc = g["C0"].copy(), 0
        for s in range(int(g["nsteps"])):
            C, sc = orc.wy_step(m, int(g["method"]), C, g["u"], g["v"], float(g["dt"]), sc)
            assert np.array_equal(C, g["C_steps"][s]), s
    finally:
        orc.set_top_jet(False)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "spobs_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_oracle_obstacle_projection_golden(orc, path):
    """From the reference's u*, v*: the Poisson block (assembly with solid
    obstacle faces, CG + SGS to 1e-7, mean removal), the correction (solid
    faces skipped) and the velocity BC reproduce its p, u, v and iterations."""
    g = np.load(path, allow_pickle=False)
    nx, ny = int(g["nx"]), int(g["ny"])
    m = orc.mesh(nx, ny, float(g["lx"]), float(g["ly"]), bc=tuple(int(b) for b in g["bc"]))
    ob = g["obs"]
    orc.set_obstacle(True, int(g["obs_type"]), *ob, perm=0.0)
    try:
        co = orc.poisson_assemble(m, g["u_star"], g["v_star"], g["rho_u"], g["rho_v"], float(g["dt"]))
        x, it, _ = orc.sp_pcg(m, 2, co, co["b"], np.zeros((ny, nx)), tol=1e-7, maxit=1000)
        p = orc.sp_finish_pressure(m, x, np.zeros((ny + 2, nx + 2)))
        u, v = orc.sp_correct(m, p, g["rho_u"], g["rho_v"], float(g["dt"]), g["u_star"], g["v_star"])
        orc.velocity_bc(m, {}, u, v)
    finally:
        orc.set_obstacle(False)
    assert it == int(g["iters"]) - int(g["iters_u"])
    assert np.array_equal(p[1:ny + 1, 1:nx + 1], g["p"][1:ny + 1, 1:nx + 1])
    assert np.array_equal(u, g["u"]) and np.array_equal(v, g["v"])
    # the obstacle changes the answer: without it the same inputs give another p
    co0 = orc.poisson_assemble(m, g["u_star"], g["v_star"], g["rho_u"], g["rho_v"], float(g["dt"]))
    assert not np.array_equal(co0["aW"], co["aW"])
