"""The parallel preconditioners of the Poisson CG (no reference counterpart):
red-black symmetric Gauss-Seidel and the multigrid V-cycle.  CG needs them to
be fixed symmetric positive definite operators, and a solve with them must
land on the same solution as the reference's Jacobi-PCG."""
import numpy as np
import pytest

from conftest import to_gpu

pytestmark = pytest.mark.gpu

CASES = [((96, 80), (4, 4, 4, 4)), ((128, 64), (3, 3, 3, 3)), ((100, 60), (1, 2, 3, 3)), ((70, 90), (2, 2, 4, 4)),
         ((40, 30), (4, 4, 3, 3))]
IDS = ["walls", "periodic", "wall-neumann-periodic", "odd-coarsening", "one-workgroup"]


def _ctx(cv, shape, bc, seed=0, with_rho=False):
    from covof_amd import workloads as wl
    nx, ny = shape
    n = max(nx, ny)
    c = wl.poisson_case(n, seed=seed, n_circles=6, rmin=n / 24, rmax=n / 6)
    C = c["C"][:ny + 2, :nx + 2].copy()
    rho = 1000.0 * C + (1.0 - C)
    rho_u = np.ones_like(rho)
    rho_v = np.ones_like(rho)
    rho_u[:, 1:] = 0.5 * (rho[:, :-1] + rho[:, 1:])
    rho_v[1:, :] = 0.5 * (rho[:-1, :] + rho[1:, :])
    if bc[0] == 3:
        rho_u[:, nx + 1] = rho_u[:, 1]
    if bc[2] == 3:
        rho_v[ny + 1, :] = rho_v[1, :]
    ctx = cv.Context(nx, ny, bc=bc)
    z = ctx.new_field()
    ctx.poisson_assemble(z, z, to_gpu(rho_u), to_gpu(rho_v), 1.0)
    return (ctx, (rho_u, rho_v)) if with_rho else ctx


@pytest.mark.parametrize("shape,bc", CASES, ids=IDS)
@pytest.mark.parametrize("pc", ["SGS_RB", "MG"])
def test_preconditioner_symmetric_positive(cv, shape, bc, pc):
    precond = getattr(cv, "PRECOND_" + pc)
    ctx = _ctx(cv, shape, bc)
    nx, ny = shape
    rng = np.random.default_rng(1)
    r1, r2 = (rng.normal(size=(ny, nx)) for _ in range(2))
    z1 = ctx.poisson_precond(precond, to_gpu(r1)).cpu().numpy()
    z2 = ctx.poisson_precond(precond, to_gpu(r2)).cpu().numpy()
    a, b = (z1 * r2).sum(), (r1 * z2).sum()
    assert abs(a - b) <= 1e-11 * np.sqrt((z1 * r1).sum() * (z2 * r2).sum()), (a, b)
    assert (z1 * r1).sum() > 0 and (z2 * r2).sum() > 0
    # deterministic: the same input gives the same bits
    assert np.array_equal(ctx.poisson_precond(precond, to_gpu(r1)).cpu().numpy(), z1)


@pytest.mark.parametrize("shape,bc", CASES, ids=IDS)
def test_mg_pcg_matches_jacobi_pcg(cv, orc, shape, bc):
    ctx = _ctx(cv, shape, bc)
    nx, ny = shape
    rng = np.random.default_rng(2)
    b = rng.normal(size=(ny, nx))
    if bc[0] != 2 and bc[1] != 2:  # singular operator: a consistent right-hand side
        b -= b.mean()
    bg = to_gpu(b)
    res = {}
    for name in ("JACOBI", "SGS_RB", "MG"):
        for prec in (cv.PRECISION_FP64, cv.PRECISION_MIXED):
            x = ctx.new_work()
            r = ctx.poisson_solve(bg, x, precond=getattr(cv, "PRECOND_" + name), tol=1e-11, max_iter=100000,
                                  precision=prec)
            assert r.converged, (name, prec, r)
            res[(name, prec)] = (x.cpu().numpy(), r.iters)
    print({k: v[1] for k, v in res.items()})
    xj = res[("JACOBI", cv.PRECISION_FP64)][0]
    xj = xj - xj.mean()
    for k, (x, _) in res.items():
        assert np.abs((x - x.mean()) - xj).max() <= 1e-8 * np.abs(xj).max(), k
    assert res[("MG", cv.PRECISION_FP64)][1] * 5 < res[("JACOBI", cv.PRECISION_FP64)][1]
    # BiCGSTAB(2) takes the same preconditioners
    x = ctx.new_work()
    r = ctx.poisson_solve(bg, x, linear_solver=cv.LINEAR_BICGSTAB2, precond=cv.PRECOND_MG, tol=1e-11,
                          max_iter=10000)
    assert r.converged, r
    xn = x.cpu().numpy()
    assert np.abs((xn - xn.mean()) - xj).max() <= 1e-7 * np.abs(xj).max()


def test_mg_mixed_2048_reaches_1e10(cv, orc):
    """The 2048^2 Poisson workload: MG-PCG to 1e-10 in FP64 and in both mixed
    forms (FP64 CG with the V-cycle in FP32, the default; the FP32 Krylov
    recurrence with FP64 reliable updates), each checked by its FP64 true
    residual on the host with the oracle's matvec.  The FP32 cycle keeps the
    FP64 solve's iteration count to within two, and the solve reports the
    true residual."""
    from covof_amd import workloads as wl
    n = 2048
    case = wl.poisson_case(n, seed=0)
    ctx = cv.Context(n, n, bc=case["bc"])
    b = ctx.poisson_assemble(*(to_gpu(case[k]) for k in ("u", "v", "rho_u", "rho_v")), case["dt"])
    out = {}
    for name, prec, form in (("fp64", cv.PRECISION_FP64, 0), ("mixed", cv.PRECISION_MIXED, 0),
                             ("mixed-fp32-krylov", cv.PRECISION_MIXED, 1)):
        ctx.set_solver_option(cv.OPT_MIXED_MG_FORM, form)
        x = ctx.new_work()
        r = ctx.poisson_solve(b, x, precond=cv.PRECOND_MG, tol=1e-10, max_iter=10000, precision=prec)
        assert r.converged, (name, r)
        out[name] = (x.cpu().numpy(), r)
    m = orc.mesh(n, n, bc=case["bc"])
    co = orc.poisson_assemble(m, case["u"], case["v"], case["rho_u"], case["rho_v"], case["dt"])
    for name, (x, r) in out.items():
        rr = np.linalg.norm(co["b"] - orc.poisson_matvec(m, co, x)) / np.linalg.norm(co["b"])
        print(f"{name}: {r.iters} MG-PCG iterations, true relres {rr:.3e} (reported {r.relres:.3e})")
        assert rr <= 1.01e-10
    rr_mixed = out["mixed"][1].relres
    x = out["mixed"][0]
    rr = np.linalg.norm(co["b"] - orc.poisson_matvec(m, co, x)) / np.linalg.norm(co["b"])
    assert abs(rr - rr_mixed) <= 1e-3 * rr
    assert out["mixed"][1].iters <= out["fp64"][1].iters + 2


def test_mg_mixed_stall_continues_in_fp64(cv, orc):
    """With tol 0 the mixed MG-PCG's FP64 residual reaches its rounding floor
    and stops improving; the solve then continues as FP64 PCG from the refined
    x until the iteration cap (NOT_CONVERGED with iters == cap) or until the
    FP64 recurrence's own breakdown guard (BREAKDOWN, which only the FP64
    loop reports).  Either way the answer sits at the FP64 floor."""
    from covof_amd import workloads as wl
    n, cap = 256, 400
    case = wl.poisson_case(n, seed=1, n_circles=12, rmin=8, rmax=32)
    ctx = cv.Context(n, n, bc=case["bc"])
    ctx.set_solver_option(cv.OPT_MIXED_MG_FORM, 1)  # the FP32 Krylov recurrence
    b = ctx.poisson_assemble(*(to_gpu(case[k]) for k in ("u", "v", "rho_u", "rho_v")), case["dt"])
    x = ctx.new_work()
    r = ctx.poisson_solve(b, x, precond=cv.PRECOND_MG, tol=0.0, max_iter=cap, precision=cv.PRECISION_MIXED)
    it32, it64 = ctx.mixed_solve_info()
    print(f"mixed MG-PCG, tol 0: status {r.status}, {r.iters} iterations ({it32} FP32 + {it64} FP64), "
          f"relres {r.relres:.3e}")
    assert r.status in (cv._lib.NOT_CONVERGED, cv._lib.BREAKDOWN), r
    assert it64 > 0 and it32 + it64 == r.iters  # the FP64 continuation ran
    if r.status == cv._lib.NOT_CONVERGED:
        assert r.iters == cap, r
    m = orc.mesh(n, n, bc=case["bc"])
    co = orc.poisson_assemble(m, case["u"], case["v"], case["rho_u"], case["rho_v"], case["dt"])
    rr = np.linalg.norm(co["b"] - orc.poisson_matvec(m, co, x.cpu().numpy())) / np.linalg.norm(co["b"])
    assert rr <= 1e-12, rr


def test_mixed_fp64_continuation_reaches_tolerance(cv, orc):
    """A stall forced early (two FP32 iterations per reliable update, a window
    of one update) hands the solve to the FP64 continuation, which ends it at
    tol 1e-10: status OK, both stages counted, and the reported relres is the
    true FP64 residual of the returned x (the oracle's matvec agrees)."""
    from covof_amd import workloads as wl
    n = 128
    case = wl.poisson_case(n, seed=2, n_circles=8, rmin=6, rmax=24)
    ctx = cv.Context(n, n, bc=case["bc"])
    ctx.set_solver_option(cv.OPT_MIXED_CYCLE, 2)
    ctx.set_solver_option(cv.OPT_MIXED_STALL_WINDOW, 1)
    b = ctx.poisson_assemble(*(to_gpu(case[k]) for k in ("u", "v", "rho_u", "rho_v")), case["dt"])
    x = ctx.new_work()
    r = ctx.poisson_solve(b, x, precond=cv.PRECOND_JACOBI, tol=1e-10, max_iter=100000, precision=cv.PRECISION_MIXED)
    it32, it64 = ctx.mixed_solve_info()
    print(f"forced stall: status {r.status}, {it32} FP32 + {it64} FP64 iterations, relres {r.relres:.3e}")
    assert r.status == cv._lib.OK, r
    assert it32 > 0 and it64 > 0 and it32 + it64 == r.iters
    m = orc.mesh(n, n, bc=case["bc"])
    co = orc.poisson_assemble(m, case["u"], case["v"], case["rho_u"], case["rho_v"], case["dt"])
    rr = np.linalg.norm(co["b"] - orc.poisson_matvec(m, co, x.cpu().numpy())) / np.linalg.norm(co["b"])
    assert rr <= 1e-10 and abs(rr - r.relres) <= 1e-3 * rr, (rr, r.relres)
    # the default window on a fresh context: no stall on this system
    ctx2 = cv.Context(n, n, bc=case["bc"])
    b2 = ctx2.poisson_assemble(*(to_gpu(case[k]) for k in ("u", "v", "rho_u", "rho_v")), case["dt"])
    r2 = ctx2.poisson_solve(b2, ctx2.new_work(), precond=cv.PRECOND_JACOBI, tol=1e-10, max_iter=100000,
                            precision=cv.PRECISION_MIXED)
    assert r2.status == cv._lib.OK and ctx2.mixed_solve_info()[1] == 0


def test_solver_options_validated(cv):
    """Out-of-range tunings are refused; a changed multigrid option takes
    effect on the next solve (the captured cycles are dropped)."""
    from covof_amd import workloads as wl
    ctx = cv.Context(64, 64, bc=(1, 1, 1, 1))
    for opt, bad in ((cv.OPT_MIXED_CYCLE, 3), (cv.OPT_MIXED_CYCLE, 1), (cv.OPT_MIXED_STALL_WINDOW, -1),
                     (cv.OPT_MIXED_LOOP, 2), (cv.OPT_MG_SWEEPS, 5), (cv.OPT_MG_OMEGA, 2.5), (cv.OPT_MG_COARSE, 0.5), (cv.OPT_MG_FUSED, 2), (cv.OPT_MIXED_MG_FORM, 3), (cv.OPT_MG_PROJECT, 2),
                     (99, 1)):
        with pytest.raises(cv.CovofError):
            ctx.set_solver_option(opt, bad)
    case = wl.poisson_case(64, seed=4, n_circles=4, rmin=4, rmax=12)
    ctx = cv.Context(64, 64, bc=case["bc"])
    b = ctx.poisson_assemble(*(to_gpu(case[k]) for k in ("u", "v", "rho_u", "rho_v")), case["dt"])
    its = []
    for sweeps in (1, 2, 1):
        ctx.set_solver_option(cv.OPT_MG_SWEEPS, sweeps)
        r = ctx.poisson_solve(b, ctx.new_work(), precond=cv.PRECOND_MG, tol=1e-10, max_iter=1000)
        assert r.converged, r
        its.append(r.iters)
    assert its[0] == its[2] and its[1] < its[0], its


FUSED_CASES = [((300, 260), (4, 4, 4, 4)), ((256, 192), (3, 3, 3, 3)), ((200, 136), (1, 2, 3, 3)),
               ((130, 514), (2, 2, 4, 4)), ((8, 2048), (3, 3, 4, 4)), ((1030, 512), (4, 4, 3, 3))]
FUSED_IDS = ["walls-odd", "periodic", "wall-neumann-periodic", "neumann-tall", "narrow-periodic-x",
             "wide-periodic-y"]


@pytest.mark.parametrize("shape,bc", FUSED_CASES, ids=FUSED_IDS)
def test_mg_fused_levels_bitwise(cv, shape, bc):
    """The fused level passes (smoothing + restriction in one kernel,
    prolongation + both smoothing half-sweeps in another) evaluate the
    five-launch V-cycle's expressions on the same operands: M r is bitwise
    the same, and so are whole FP64 and mixed-precision MG-PCG solves (both
    mixed forms: FP64 CG around FP32 V-cycles, where the first level's passes
    round r and widen z, and the FP32 Krylov recurrence): status, iteration
    count and x.  (The narrow
    periodic strip converges slowly -- its coarsest level is a 2 x 512 line
    that the bottom's sweeps do not solve -- which does not matter here.)"""
    ctx = _ctx(cv, shape, bc)
    nx, ny = shape
    rng = np.random.default_rng(5)
    r = to_gpu(rng.normal(size=(ny, nx)))
    z = {}
    for fused in (1, 0):
        ctx.set_solver_option(cv.OPT_MG_FUSED, fused)
        z[fused] = ctx.poisson_precond(cv.PRECOND_MG, r).cpu().numpy()
    assert np.array_equal(z[0], z[1])
    b = rng.normal(size=(ny, nx))
    if bc[0] != 2 and bc[1] != 2:
        b -= b.mean()
    bg = to_gpu(b)
    out = {}
    runs = ((cv.PRECISION_FP64, 0), (cv.PRECISION_MIXED, 0), (cv.PRECISION_MIXED, 1))
    for fused in (1, 0):
        ctx.set_solver_option(cv.OPT_MG_FUSED, fused)
        for prec, form in runs:
            ctx.set_solver_option(cv.OPT_MIXED_MG_FORM, form)
            x = ctx.new_work()
            res = ctx.poisson_solve(bg, x, precond=cv.PRECOND_MG, tol=1e-10, max_iter=400, precision=prec)
            out[(fused, prec, form)] = (res.status, res.iters, x.cpu().numpy())
    for prec, form in runs:
        assert out[(1, prec, form)][:2] == out[(0, prec, form)][:2]
        assert np.array_equal(out[(1, prec, form)][2], out[(0, prec, form)][2])


PIN_CASES = [((96, 80), (4, 4, 4, 4)), ((300, 260), (4, 4, 4, 4)), ((256, 192), (3, 3, 3, 3)),
             ((200, 136), (1, 2, 3, 3)), ((70, 90), (2, 2, 4, 4)), ((8, 2048), (3, 3, 4, 4)), ((40, 30), (4, 4, 3, 3)),
             ((206, 128), (1, 1, 3, 3))]
PIN_IDS = ["walls-small", "walls-odd", "periodic", "wall-neumann-periodic", "odd-coarsening", "narrow-periodic-x",
           "one-workgroup", "odd-periodic-y"]


@pytest.mark.parametrize("shape,bc", PIN_CASES, ids=PIN_IDS)
def test_mg_vcycle_bitwise_vs_numpy(cv, orc, shape, bc):
    """M r of PRECOND_MG (FP64) bit for bit against the numpy restatement of
    the cycle (tests/mg_numpy.py) on the oracle's assembled operator: the
    hierarchy, the fused multi-workgroup levels, the single-workgroup bottom
    and its wave-only coarse levels, and the mean-free coarsest solve of a
    singular operator (on and off)."""
    import mg_numpy as mgn
    nx, ny = shape
    ctx, (rho_u, rho_v) = _ctx(cv, shape, bc, with_rho=True)
    m = orc.mesh(nx, ny, bc=bc)
    zf = np.zeros((ny + 2, nx + 2))
    co = orc.poisson_assemble(m, zf, zf, rho_u, rho_v, 1.0)
    kx, ky = mgn.faces_from_assembly(co)
    perx, pery = int(bc[0] == 3), int(bc[2] == 3)
    levels = mgn.hierarchy(kx, ky, co["aP"], perx, pery, 1.0e-12 * (1.0 / nx) * (1.0 / ny))
    singular = 2 not in bc  # no Dirichlet pressure side
    rng = np.random.default_rng(6)
    for trial in range(2):
        r = rng.normal(size=(ny, nx))
        for project in ((1, 0) if singular else (1,)):
            ctx.set_solver_option(cv.OPT_MG_PROJECT, project)
            zn = mgn.vcycle(levels, r, perx, pery, project=bool(project and singular))
            for fused in (1, 0):
                ctx.set_solver_option(cv.OPT_MG_FUSED, fused)
                zg = ctx.poisson_precond(cv.PRECOND_MG, to_gpu(r)).cpu().numpy()
                bad = np.argwhere(zg != zn)
                assert bad.size == 0, (project, fused, len(bad), bad[:4], np.abs(zg - zn).max())
