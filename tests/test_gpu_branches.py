"""Branches of the hot path's steps, each against the compiled reference's
fixtures (tests/golden/props_*, jet_*, spobs_*) and the oracle:

* update_physical_properties with SVOF smoothing on: rho, mu and mu_node from
  C_s (src/module_interface_vof.f90:540-545, :596-600);
* the JET_INLET top profile of apply_vof_bc inside the Weymouth-Yue transport
  (src/module_mpi_comm.f90:664-676);
* solid-obstacle faces in the SP Poisson assembly and the velocity correction
  (src/module_simulation.f90:2491-2528, :2596-2632).
"""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, to_gpu

pytestmark = pytest.mark.gpu

METHOD = {0: "PLIC", 1: "ELVIRA"}


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "props_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_properties_golden(cv, path):
    g = np.load(path, allow_pickle=False)
    nx, ny = int(g["nx"]), int(g["ny"])
    ctx = cv.Context(nx, ny, float(g["lx"]), float(g["ly"]), bc=tuple(int(b) for b in g["bc"]))
    C = to_gpu(g["C0"])
    Cs = to_gpu(g["C_s"]) if int(g["svof"]) else None
    props = ctx.update_physical_properties(C, float(g["rho1"]), float(g["rho2"]), float(g["mu1"]), float(g["mu2"]),
                                           C_s=Cs)
    assert np.array_equal(C.cpu().numpy(), g["C"])
    sl = {"rho": np.s_[:, :], "mu": np.s_[:, :], "rho_u": np.s_[1:ny + 1, 1:], "rho_v": np.s_[1:, 1:nx + 1],
          "mu_node": np.s_[1:, 1:]}
    for k, s in sl.items():
        got = props[k].cpu().numpy()
        assert np.array_equal(got[s], g[k][s]), f"{k}: {_rel(got[s], g[k][s]):.3e}"


def test_properties_svof_vs_oracle(cv, orc):
    """The SVOF branch on a periodic-x / wall-y grid against the oracle, and
    that it differs from the C branch."""
    from covof_amd import workloads as wl
    nx, ny = 70, 52
    C = wl.ellipses(nx, ny, 9, 3.0, 12.0, seed=81)
    C[:, 0], C[:, nx + 1] = C[:, nx], C[:, 1]
    C[0, :], C[ny + 1, :] = C[1, :], C[ny, :]
    rng = np.random.default_rng(82)
    Cs = np.clip(C + 0.1 * rng.normal(size=C.shape), -0.1, 1.1)
    m = orc.mesh(nx, ny, bc=(3, 3, 4, 1))
    Co = C.copy()
    ref = orc.update_properties_cs(m, Co, Cs, 1000.0, 1.0, 1.0, 0.001)
    ctx = cv.Context(nx, ny, bc=(3, 3, 4, 1))
    Cg = to_gpu(C)
    props = ctx.update_physical_properties(Cg, 1000.0, 1.0, 1.0, 0.001, props={k: to_gpu(np.ones_like(C)) for k in
                                                                              ("rho", "rho_u", "rho_v", "mu",
                                                                               "mu_node")}, C_s=to_gpu(Cs))
    for k, a in zip(("rho", "rho_u", "rho_v", "mu", "mu_node"), ref):
        assert np.array_equal(props[k].cpu().numpy(), a), k
    plain = ctx.update_physical_properties(to_gpu(C), 1000.0, 1.0, 1.0, 0.001)
    assert not np.array_equal(plain["mu_node"].cpu().numpy(), ref[4])


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "jet_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_jet_inlet_golden(cv, path):
    g = np.load(path, allow_pickle=False)
    nx, ny = int(g["nx"]), int(g["ny"])
    ctx = cv.Context(nx, ny, float(g["lx"]), float(g["ly"]), bc=tuple(int(b) for b in g["bc"]))
    ctx.set_vof_top_profile("JET_INLET", float(g["jet_center"]), float(g["jet_width"]))
    C, u, v = to_gpu(g["C0"]), to_gpu(g["u"]), to_gpu(g["v"])
    for s in range(int(g["nsteps"])):
        ctx.advect_vof(C, u, v, float(g["dt"]), reconstruction_method=METHOD[int(g["method"])])
        got = C.cpu().numpy()
        assert np.array_equal(got, g["C_steps"][s]), (s, _rel(got, g["C_steps"][s]))
    # the profile is context state: UNIFORM restores the zero-gradient top halo
    ctx.set_vof_top_profile("UNIFORM")
    ctx.advect_vof(C, u, v, float(g["dt"]), reconstruction_method=METHOD[int(g["method"])])
    Cn = C.cpu().numpy()
    assert np.array_equal(Cn[ny + 1, 1:nx + 1], Cn[ny, 1:nx + 1])


@pytest.mark.parametrize("order", ["reference", "tree"])
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "spobs_*.npz"))),
                         ids=lambda p: os.path.basename(p))
def test_obstacle_projection_golden(cv, path, order):
    """From the reference's u*, v*: the GPU Poisson block (assembly with the
    obstacle's solid faces, CG + SGS to 1e-7, mean removal), the correction
    and the velocity BC.  The reference's summation order gives its p, u, v
    bit for bit and its iteration count; tree sums agree within 1e-6."""
    g = np.load(path, allow_pickle=False)
    nx, ny = int(g["nx"]), int(g["ny"])
    ctx = cv.Context(nx, ny, float(g["lx"]), float(g["ly"]), bc=tuple(int(b) for b in g["bc"]))
    ob = [float(a) for a in g["obs"]]
    ctx.set_obstacle(int(g["obs_type"]), *ob[:4], xc=ob[4], yc=ob[5], radius=ob[6], perm=0.0)
    if order == "reference":
        ctx.set_reduction_order(cv.REDUCTION_REFERENCE)
    dt = float(g["dt"])
    p = ctx.new_field()
    u, v, ru, rv = (to_gpu(g[k]) for k in ("u_star", "v_star", "rho_u", "rho_v"))
    res = ctx.solve_standard_projection_poisson(u, v, ru, rv, p, dt)
    ctx.correct_velocity(p, ru, rv, dt, u, v)
    ctx.apply_velocity_bc(cv.Context.momentum_cfg(), u, v)
    it_ref = int(g["iters"]) - int(g["iters_u"])
    pg, ug, vg = p.cpu().numpy(), u.cpu().numpy(), v.cpu().numpy()
    I = np.s_[1:ny + 1, 1:nx + 1]
    print(f"[{os.path.basename(path)} {order}] iters {res.iters} (reference {it_ref}), p {_rel(pg[I], g['p'][I]):.2e}")
    if order == "reference":
        assert res.iters == it_ref
        assert np.array_equal(pg[I], g["p"][I])
        assert np.array_equal(ug, g["u"]) and np.array_equal(vg, g["v"])
    else:
        assert abs(res.iters - it_ref) <= 1
        assert _rel(pg[I], g["p"][I]) < 1e-6 and _rel(ug, g["u"]) < 1e-6 and _rel(vg, g["v"]) < 1e-6
    # the solid faces carry no correction and the momentum block refuses the obstacle
    with pytest.raises(cv.CovofError):
        ctx.momentum_set_fields(cv.Context.momentum_cfg(), {k: ctx.new_field(1.0) for k in
                                                            cv.Context.MOMENTUM_FIELDS}, dt)


def test_obstacle_assembly_vs_oracle(cv, orc):
    """The assembled operator with a cylinder on a periodic-x channel: its
    products bit for bit the oracle's, and the obstacle's faces are walls."""
    from covof_amd import workloads as wl
    nx, ny = 96, 64
    case = wl.poisson_case(nx, seed=3, n_circles=5, rmin=4, rmax=10)
    bc = (3, 3, 1, 1)
    m = orc.mesh(nx, nx, bc=bc)
    ctx = cv.Context(nx, nx, bc=bc)
    ctx.set_obstacle(2, xc=0.5, yc=0.45, radius=0.2, perm=0.0)
    orc.set_obstacle(True, 2, xc=0.5, yc=0.45, radius=0.2, perm=0.0)
    try:
        T = {k: to_gpu(case[k]) for k in ("u", "v", "rho_u", "rho_v")}
        b = ctx.poisson_assemble(T["u"], T["v"], T["rho_u"], T["rho_v"], case["dt"])
        co = orc.poisson_assemble(m, case["u"], case["v"], case["rho_u"], case["rho_v"], case["dt"])
        assert np.array_equal(b.cpu().numpy(), co["b"])
        x = np.random.default_rng(4).normal(size=(nx, nx))
        assert np.array_equal(ctx.poisson_matvec(to_gpu(x)).cpu().numpy(), orc.poisson_matvec(m, co, x))
        assert (co["aW"] == 0).sum() > nx  # faces inside the disc are walls
    finally:
        orc.set_obstacle(False)
