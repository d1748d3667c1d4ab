"""The Fortran side of the boundary: covof_driver (ISO_C_BINDING module
covof_hip) calls the host-array entry points the way the reference's time loop
calls advect_vof and the SP Poisson block; its results must equal the oracle's
(VOF, bit-exact) and the Python path's over the same C-ABI (bit-exact)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, to_gpu

DRIVER = os.path.join(ROOT, "covof_2d-fortran_amd", "fortran", "covof_driver")


def _read(path, narrays):
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(24), dtype=np.int32)
        dt, relres = np.frombuffer(f.read(16), dtype=np.float64)
        n = int(hdr[1])
        rest = np.frombuffer(f.read(), dtype=np.float64)
    assert rest.size == narrays * (n + 2) * (n + 2)
    arrays = [a.reshape(n + 2, n + 2).copy() for a in np.split(rest, narrays)]
    return hdr, float(dt), float(relres), arrays


def test_driver_links_and_prints_usage():
    """CPU: the driver starts (libcovof_hip.so resolves through its runpath)."""
    assert os.path.exists(DRIVER), "build the Fortran driver with __graft_entry__.build()"
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stdout


@pytest.mark.gpu
def test_fortran_vof_advect_matches_oracle(tmp_path, orc):
    n, steps = 64, 12
    out = tmp_path / "vof.bin"
    r = subprocess.run([DRIVER, "vof", str(n), str(steps), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    hdr, dt, _, (C0, u, v, C) = _read(out, 4)
    assert hdr[0] == 1 and hdr[4] == steps and hdr[5] == 0
    m = orc.mesh(n, n, bc=(3, 3, 3, 3))
    Co, sc = C0.copy(), 0
    for _ in range(steps):
        Co, sc = orc.wy_step(m, orc.ELVIRA, Co, u, v, dt, sc)
    assert sc == steps
    assert np.array_equal(C, Co)
    assert abs(C[1:n + 1, 1:n + 1].sum() - C0[1:n + 1, 1:n + 1].sum()) <= 1e-12 * C0[1:n + 1, 1:n + 1].sum()


@pytest.mark.gpu
def test_fortran_sp_poisson_matches_python_and_oracle(tmp_path, cv, orc):
    n = 64
    out = tmp_path / "sp.bin"
    r = subprocess.run([DRIVER, "sp", str(n), "1", str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    hdr, dt, relres, (us, vs, rho_u, rho_v, p, u, v) = _read(out, 7)
    assert hdr[0] == 2 and hdr[5] == 0
    iters = int(hdr[4])
    # the same C-ABI from Python on device arrays: identical bits
    ctx = cv.Context(n, n, bc=(cv.BC_FREESLIP,) * 4)
    pg = to_gpu(np.zeros_like(p))
    res = ctx.solve_standard_projection_poisson(to_gpu(us), to_gpu(vs), to_gpu(rho_u), to_gpu(rho_v), pg, dt)
    ug, vg = to_gpu(us), to_gpu(vs)
    ctx.correct_velocity(pg, to_gpu(rho_u), to_gpu(rho_v), dt, ug, vg)
    assert res.iters == iters and res.relres == relres
    assert np.array_equal(pg.cpu().numpy(), p)
    assert np.array_equal(ug.cpu().numpy(), u) and np.array_equal(vg.cpu().numpy(), v)
    # and the oracle's SP loop (CG + SGS, 1e-7, 1000): iterations +-1, p to 1e-6
    m = orc.mesh(n, n, bc=(4, 4, 4, 4))
    co = orc.poisson_assemble(m, us, vs, rho_u, rho_v, dt)
    xo, it, _ = orc.sp_pcg(m, 2, co, co["b"], np.zeros((n, n)), tol=1e-7, maxit=1000)
    po = orc.sp_finish_pressure(m, xo, np.zeros_like(p))
    assert abs(iters - it) <= 1, (iters, it)
    scale = np.abs(po[1:n + 1, 1:n + 1]).max()
    assert np.abs(p[1:n + 1, 1:n + 1] - po[1:n + 1, 1:n + 1]).max() / scale < 1e-6


@pytest.mark.gpu
def test_fortran_curvature_csf_matches_oracle(tmp_path, orc):
    """The tail of advect_vof from Fortran host arrays: SVOF and diffusion
    smoothing, curvature and CSF forces, bit-exact against the oracle."""
    n, steps = 64, 4
    out = tmp_path / "csf.bin"
    r = subprocess.run([DRIVER, "csf", str(n), str(steps), str(out)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    hdr, _, _, (C, Cs, Ct, curv, fx, fy) = _read(out, 6)
    assert hdr[0] == 3 and hdr[5] == 0
    m = orc.mesh(n, n, bc=(4, 4, 4, 4))
    assert np.array_equal(Cs, orc.svof_smooth(m, C, 1.5))
    Cto = orc.smooth_diffusion(m, C, steps)
    assert np.array_equal(Ct, Cto)
    co, fxo, fyo = orc.curvature_csf(m, Cto, C, 0.07)
    assert np.array_equal(curv, co) and np.array_equal(fx, fxo) and np.array_equal(fy, fyo)
    assert np.abs(curv).max() > 0.0


@pytest.mark.gpu
def test_fortran_sp_predictor_matches_python_and_oracle(tmp_path, cv, orc):
    """covof_sp_predictor from Fortran (a derived type bind(C) for the
    options) gives the Python device path's u*, v* bit for bit, and the
    oracle's within the solve's tolerance."""
    n = 48
    out = tmp_path / "pred.bin"
    r = subprocess.run([DRIVER, "pred", str(n), "1", str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    hdr, dt, relres, (u0, v0, um, vm, rho_u, rho_v, mu, mun, us, vs) = _read(out, 10)
    assert hdr[0] == 4 and hdr[5] == 0
    z = np.zeros_like(u0)
    cfg = cv.Context.momentum_cfg(time_scheme=2, has_nm1=True, convective_scheme=3, gy=-0.98, bc_u=(0, 0, 0, 1.0))
    ctx = cv.Context(n, n, bc=(1, 1, 1, 1))
    F = {"u": to_gpu(u0), "v": to_gpu(v0), "u_old": to_gpu(u0), "v_old": to_gpu(v0), "u_nm1": to_gpu(um),
         "v_nm1": to_gpu(vm), "rho_u": to_gpu(rho_u), "rho_v": to_gpu(rho_v), "mu": to_gpu(mu),
         "mu_node": to_gpu(mun), "fx_sigma": to_gpu(z), "fy_sigma": to_gpu(z)}
    res = ctx.solve_standard_projection_predictor(cfg, F, dt, 1, 1, 1e-10, 300)
    assert res.converged and res.iters == hdr[4]
    assert np.array_equal(F["u"].cpu().numpy(), us) and np.array_equal(F["v"].cpu().numpy(), vs)
    o = orc.Momentum(orc.mesh(n, n, bc=(1, 1, 1, 1)),
                     dict(u=u0, v=v0, u_old=u0, v_old=v0, u_nm1=um, v_nm1=vm, rho_u=rho_u, rho_v=rho_v, mu=mu,
                          mu_node=mun, fx=z, fy=z),
                     dict(time_scheme=2, has_nm1=1, conv_scheme=3, gy=-0.98, bc_u=(0, 0, 0, 1.0)), dt)
    uo, vo, _, _ = o.predictor(1, 1, 1e-10, 300)
    scale = max(np.abs(uo).max(), np.abs(vo).max())
    assert np.abs(us - uo).max() < 1e-8 * scale and np.abs(vs - vo).max() < 1e-8 * scale


@pytest.mark.gpu
def test_fortran_monolithic_step_matches_python(tmp_path, cv):
    """covof_monolithic_step from Fortran (solver_type MONOLITHIC in the
    bind(C) options, BiCGSTAB(2) + block-triangular Schur-SSOR) gives the
    Python device path's u, v, p bit for bit."""
    n = 40
    out = tmp_path / "mono.bin"
    r = subprocess.run([DRIVER, "mono", str(n), "1", str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    hdr, dt, relres, (u0, v0, um, vm, rho_u, rho_v, mu, mun, us, vs, ps) = _read(out, 11)
    assert hdr[0] == 5 and hdr[5] == 0
    z = np.zeros_like(u0)
    cfg = cv.Context.momentum_cfg(time_scheme=2, has_nm1=True, convective_scheme=3, gy=-0.98, bc_u=(0, 0, 0, 1.0),
                                  solver_type=cv.SOLVER_MONOLITHIC)
    ctx = cv.Context(n, n, bc=(1, 1, 1, 1))
    F = {"u": to_gpu(u0), "v": to_gpu(v0), "u_old": to_gpu(u0), "v_old": to_gpu(v0), "u_nm1": to_gpu(um),
         "v_nm1": to_gpu(vm), "rho_u": to_gpu(rho_u), "rho_v": to_gpu(rho_v), "mu": to_gpu(mu),
         "mu_node": to_gpu(mun), "fx_sigma": to_gpu(z), "fy_sigma": to_gpu(z)}
    p = to_gpu(z)
    res = ctx.solve_monolithic_step(cfg, F, p, dt, 1, 4, 1e-10, 600)
    assert res.converged and res.iters == hdr[4]
    for got, ref in ((F["u"], us), (F["v"], vs), (p, ps)):
        assert np.array_equal(got.cpu().numpy(), ref)


@pytest.mark.gpu
def test_fortran_precond_hook_body_matches_python(tmp_path, cv):
    """The PRECOND_CUSTOM hook's body (covof_precond_apply: the fields bound
    from host arrays with covof_momentum_set_fields, dst = M^-1 src through
    covof_momentum_precond) gives the device path's SSOR application bit for
    bit; a size that is not the bound system's is refused and leaves the
    hook's fallback dst = src."""
    import ctypes as ct
    from covof_amd import _lib as L
    n = 40
    out = tmp_path / "hook.bin"
    r = subprocess.run([DRIVER, "hook", str(n), "1", str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    with open(out, "rb") as f:
        hdr = np.frombuffer(f.read(24), dtype=np.int32)
        dt = float(np.frombuffer(f.read(16), dtype=np.float64)[0])
        rest = np.frombuffer(f.read(), dtype=np.float64)
    assert hdr[0] == 6 and hdr[5] == 0
    nm, m = int(hdr[4]), (n + 2) * (n + 2)
    assert rest.size == 8 * m + 2 * nm
    u0, v0, um, vm, rho_u, rho_v, mu, mun = [a.reshape(n + 2, n + 2).copy() for a in np.split(rest[:8 * m], 8)]
    src, dst = rest[8 * m:8 * m + nm].copy(), rest[8 * m + nm:].copy()
    z = np.zeros_like(u0)
    cfg = cv.Context.momentum_cfg(time_scheme=2, has_nm1=True, convective_scheme=3, gy=-0.98, bc_u=(0, 0, 0, 1.0))
    ctx = cv.Context(n, n, bc=(1, 1, 1, 1))
    F = {"u": to_gpu(u0), "v": to_gpu(v0), "u_old": to_gpu(u0), "v_old": to_gpu(v0), "u_nm1": to_gpu(um),
         "v_nm1": to_gpu(vm), "rho_u": to_gpu(rho_u), "rho_v": to_gpu(rho_v), "mu": to_gpu(mu),
         "mu_node": to_gpu(mun), "fx_sigma": to_gpu(z), "fy_sigma": to_gpu(z)}
    ctx.momentum_set_fields(cfg, F, dt)
    assert ctx.momentum_size()[0] == nm
    want = ctx.apply_preconditioner(to_gpu(src), precond=cv.PRECOND_SSOR).cpu().numpy()
    assert np.array_equal(dst, want)
    assert not np.array_equal(dst, src)
    # the host entry refuses a vector that is not the bound system's size
    lib = L.load()
    a = np.ascontiguousarray(src)
    b = np.zeros_like(a)
    dp = ct.POINTER(ct.c_double)
    st = lib.covof_momentum_precond(ctx._h, cv.PRECOND_SSOR, nm - 1, a.ctypes.data_as(dp), b.ctypes.data_as(dp))
    assert st == L.BAD_ARGUMENT
