"""ctypes front-end of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, as the checker (or the timed CPU baseline).  The product package never
imports it.

Array conventions: a halo field of the reference, Fortran ``F(0:nx+1, 0:ny+1)``
column-major with ``i`` fastest, is a C-contiguous numpy array of shape
``(ny+2, nx+2)`` indexed ``F[j, i]``; a Poisson work array ``A(1:nx, 1:ny)`` is
shape ``(ny, nx)``.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC, BC_FREESLIP = 1, 2, 3, 4
PLIC, ELVIRA = 0, 1


class MomCfg(ct.Structure):
    _fields_ = [("time_scheme", ct.c_int), ("has_nm1", ct.c_int), ("enable_convection", ct.c_int),
                ("conv_scheme", ct.c_int), ("conv_lin", ct.c_int), ("gx", ct.c_double), ("gy", ct.c_double),
                ("bc_u", ct.c_double * 4), ("bc_v", ct.c_double * 4), ("solver_type", ct.c_int)]


class MomBlocks(ct.Structure):
    _fields_ = [(k, ct.POINTER(ct.c_double)) for k in
                ("diag", "uW", "uE", "uS", "uN", "uP", "vW", "vE", "vS", "vN", "vP", "sW", "sE", "sS", "sN", "sP")]


class MomFields(ct.Structure):
    _fields_ = [(k, ct.POINTER(ct.c_double)) for k in
                ("u", "v", "u_old", "v_old", "u_nm1", "v_nm1", "rho_u", "rho_v", "mu", "mu_node", "fx", "fy")]


class Mesh(ct.Structure):
    _fields_ = [("nx", ct.c_int), ("ny", ct.c_int), ("lx", ct.c_double), ("ly", ct.c_double),
                ("dx", ct.c_double), ("dy", ct.c_double), ("bc", ct.c_int * 4)]


_lib = None
_D = ct.POINTER(ct.c_double)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ct.CDLL(_LIB)
        M = ct.POINTER(Mesh)
        d, i, ip = ct.c_double, ct.c_int, ct.POINTER(ct.c_int)
        sig = {
            "orc_mesh_init": (None, [M, i, i, d, d, ct.POINTER(ct.c_int)]),
            "orc_cut_box_area": (d, [d] * 7),
            "orc_compute_alpha": (d, [d] * 7),
            "orc_reconstruct": (None, [M, i, _D, _D, _D, _D]),
            "orc_reconstruct_interface": (None, [M, i, _D, _D, _D, _D]),
            "orc_exchange_halos_scalar": (None, [M, _D]),
            "orc_apply_vof_bc": (None, [M, _D]),
            "orc_wy_step": (None, [M, i, _D, _D, _D, d, ip]),
            "orc_update_properties": (None, [M, _D, d, d, d, d, _D, _D, _D, _D, _D]),
            "orc_update_properties_cs": (None, [M, _D, _D, d, d, d, d, _D, _D, _D, _D, _D]),
            "orc_set_obstacle": (None, [i, i] + [d] * 8),
            "orc_set_top_jet": (None, [i, d, d]),
            "orc_svof_smooth": (None, [M, _D, d, _D]),
            "orc_smooth_diffusion": (None, [M, _D, i, _D]),
            "orc_curvature_csf": (None, [M, _D, _D, i, d, _D, _D, _D]),
            "orc_curvature_csf2": (None, [M, _D, _D, i, d, i, _D, _D, _D]),
            "orc_poisson_assemble": (None, [M, _D, _D, _D, _D, d] + [_D] * 6),
            "orc_poisson_matvec": (None, [M] + [_D] * 7),
            "orc_poisson_sgs": (None, [M] + [_D] * 7),
            "orc_poisson_precond": (None, [M, i] + [_D] * 7),
            "orc_sp_pcg": (i, [M, i] + [_D] * 7 + [d, i, _D]),
            "orc_bicgstab2": (i, [M, i] + [_D] * 7 + [d, i, _D]),
            "orc_sp_finish_pressure": (None, [M, _D, _D]),
            "orc_sp_correct": (None, [M, _D, _D, _D, d, _D, _D]),
            "orc_dot": (d, [_D, _D, ct.c_long]),
            "orc_bench_jacobi_pcg": (d, [M] + [_D] * 6 + [i, i]),
            "orc_bench_jacobi_pcg_solve": (d, [M] + [_D] * 6 + [d, i, i, ct.POINTER(ct.c_int), _D]),
            "orc_set_threads": (None, [i]),
            "orc_history": (None, [M] + [_D] * 7),
            "orc_velocity_bc": (None, [M, ct.POINTER(MomCfg), _D, _D, i]),
            "orc_mom_sizes": (None, [M, ip, ip]),
            "orc_mom_pack": (None, [M, _D, _D, _D]),
            "orc_mom_unpack": (None, [M, ct.POINTER(MomCfg), _D, _D, _D, i]),
            "orc_mom_matvec": (None, [M, ct.POINTER(MomCfg), ct.POINTER(MomFields), d, _D, _D, _D, _D]),
            "orc_mom_diag": (None, [M, ct.POINTER(MomCfg), ct.POINTER(MomFields), d, _D]),
            "orc_mom_rhs": (None, [M, ct.POINTER(MomCfg), ct.POINTER(MomFields), d, _D]),
            "orc_mom_solve": (i, [M, ct.POINTER(MomCfg), ct.POINTER(MomFields), d, _D, _D, _D, i, i, d, i,
                                  _D, _D, _D]),
            "orc_sp_predictor": (i, [M, ct.POINTER(MomCfg), ct.POINTER(MomFields), d, _D, _D, i, i, d, i,
                                     _D, _D, _D]),
            "orc_mono_pack": (None, [M, _D, _D, _D, _D]),
            "orc_mono_matvec": (None, [M, ct.POINTER(MomCfg), ct.POINTER(MomFields), d, _D, _D, _D, _D, _D]),
            "orc_mom_blocks": (None, [M, ct.POINTER(MomCfg), ct.POINTER(MomFields), d, ct.POINTER(MomBlocks)]),
            "orc_mom_precond": (None, [M, ct.POINTER(MomCfg), i, ct.POINTER(MomBlocks), _D, _D]),
            "orc_mom_solve2": (i, [M, ct.POINTER(MomCfg), ct.POINTER(MomFields), d, _D, _D, _D, i, i, d, i, _D, _D,
                                   _D]),
            "orc_mono_step": (i, [M, ct.POINTER(MomCfg), ct.POINTER(MomFields), d, _D, _D, _D, i, i, d, i, _D, _D,
                                  _D, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def mesh(nx: int, ny: int, lx: float = 1.0, ly: float = 1.0, bc=(BC_PERIODIC,) * 4) -> Mesh:
    m = Mesh()
    arr = (ct.c_int * 4)(*bc)
    lib().orc_mesh_init(ct.byref(m), nx, ny, lx, ly, arr)
    return m


def cut_box_area(xmin, xmax, ymin, ymax, nx, ny, alpha) -> float:
    return lib().orc_cut_box_area(xmin, xmax, ymin, ymax, nx, ny, alpha)


def compute_alpha(xmin, xmax, ymin, ymax, nx, ny, c) -> float:
    return lib().orc_compute_alpha(xmin, xmax, ymin, ymax, nx, ny, c)


def reconstruct_interface(m: Mesh, method: int, C: np.ndarray):
    shp = (m.ny + 2, m.nx + 2)
    C = np.ascontiguousarray(C, dtype=np.float64)
    sx, sy, sa = (np.zeros(shp) for _ in range(3))
    lib().orc_reconstruct_interface(ct.byref(m), method, _p(C), _p(sx), _p(sy), _p(sa))
    return sx, sy, sa


def wy_step(m: Mesh, method: int, C: np.ndarray, u: np.ndarray, v: np.ndarray, dt: float,
            step_count: int):
    """advect_vof(WEYMOUTH_YUE) in place on a copy; returns (C_new, step_count_new)."""
    C = np.array(C, dtype=np.float64, order="C", copy=True)
    sc = ct.c_int(step_count)
    lib().orc_wy_step(ct.byref(m), method, _p(C), _p(np.ascontiguousarray(u)),
                      _p(np.ascontiguousarray(v)), dt, ct.byref(sc))
    return C, sc.value


def svof_smooth(m: Mesh, C: np.ndarray, width: float) -> np.ndarray:
    """smooth_svof_helmholtz: C_s (halo field) from C."""
    Cs = np.empty((m.ny + 2, m.nx + 2))
    lib().orc_svof_smooth(ct.byref(m), _p(np.ascontiguousarray(C, dtype=np.float64)), width, _p(Cs))
    return Cs


def smooth_diffusion(m: Mesh, C: np.ndarray, steps: int) -> np.ndarray:
    """vof_smooth_diffusion: C_tilde (halo field) from C."""
    Ct = np.empty((m.ny + 2, m.nx + 2))
    lib().orc_smooth_diffusion(ct.byref(m), _p(np.ascontiguousarray(C, dtype=np.float64)), steps, _p(Ct))
    return Ct


def curvature_csf(m: Mesh, G: np.ndarray, C: np.ndarray, sigma: float, enable: int = 1, fx0=None, fy0=None,
                  method: int = 0):
    """compute_curvature_and_csf: (curvature, fx_sigma, fy_sigma); normals from
    G, force gradient from C.  fx0 / fy0: the force arrays' previous contents
    (the entries outside the face ranges keep them; zeros by default).
    method 1: height-function curvature where the heights of C bracket the
    interface (no reference counterpart)."""
    shp = (m.ny + 2, m.nx + 2)
    curv = np.zeros(shp)
    fx = np.zeros(shp) if fx0 is None else np.array(fx0, dtype=np.float64, copy=True)
    fy = np.zeros(shp) if fy0 is None else np.array(fy0, dtype=np.float64, copy=True)
    lib().orc_curvature_csf2(ct.byref(m), _p(np.ascontiguousarray(G, dtype=np.float64)),
                             _p(np.ascontiguousarray(C, dtype=np.float64)), enable, ct.c_double(sigma), int(method),
                             _p(curv), _p(fx), _p(fy))
    return curv, fx, fy


def update_properties(m: Mesh, C: np.ndarray, rho1, rho2, mu1, mu2, init=1.0):
    shp = (m.ny + 2, m.nx + 2)
    out = [np.full(shp, init) for _ in range(5)]
    lib().orc_update_properties(ct.byref(m), _p(C), rho1, rho2, mu1, mu2, *[_p(a) for a in out])
    return out  # rho, rho_u, rho_v, mu, mu_node


def update_properties_cs(m: Mesh, C: np.ndarray, Cs, rho1, rho2, mu1, mu2, init=1.0):
    """update_physical_properties with the SVOF branch when Cs is given (C's
    physical corners are fixed in place)."""
    shp = (m.ny + 2, m.nx + 2)
    out = [np.full(shp, init) for _ in range(5)]
    lib().orc_update_properties_cs(ct.byref(m), _p(C), None if Cs is None else _p(np.ascontiguousarray(Cs)),
                                   rho1, rho2, mu1, mu2, *[_p(a) for a in out])
    return out  # rho, rho_u, rho_v, mu, mu_node


def set_obstacle(enabled=False, type=0, x0=0.0, y0=0.0, x1=0.0, y1=0.0, xc=0.0, yc=0.0, radius=0.0, perm=1e-40):
    """The solid obstacle the Poisson assembly and correction see (config_t's obs_*)."""
    lib().orc_set_obstacle(int(bool(enabled)), int(type), x0, y0, x1, y1, xc, yc, radius, perm)


def set_top_jet(on=False, center=0.5, width=0.08):
    """The JET_INLET top profile of apply_vof_bc (bc_top%profile, jet_center, jet_width)."""
    lib().orc_set_top_jet(int(bool(on)), center, width)


def poisson_assemble(m: Mesh, u, v, rho_u, rho_v, dt):
    shp = (m.ny, m.nx)
    aW, aE, aS, aN, aP, b = (np.zeros(shp) for _ in range(6))
    lib().orc_poisson_assemble(ct.byref(m), _p(u), _p(v), _p(rho_u), _p(rho_v), dt,
                               _p(aW), _p(aE), _p(aS), _p(aN), _p(aP), _p(b))
    return dict(aW=aW, aE=aE, aS=aS, aN=aN, aP=aP, b=b)


def _co(co):
    return [_p(co[k]) for k in ("aP", "aW", "aE", "aS", "aN")]


def poisson_matvec(m: Mesh, co, x):
    y = np.zeros_like(co["aP"])
    lib().orc_poisson_matvec(ct.byref(m), *_co(co), _p(np.ascontiguousarray(x)), _p(y))
    return y


def poisson_sgs(m: Mesh, co, r):
    z = np.zeros_like(co["aP"])
    lib().orc_poisson_sgs(ct.byref(m), *_co(co), _p(np.ascontiguousarray(r)), _p(z))
    return z


def poisson_precond(m: Mesh, precond: int, co, r):
    z = np.zeros_like(co["aP"])
    lib().orc_poisson_precond(ct.byref(m), precond, *_co(co), _p(np.ascontiguousarray(r)), _p(z))
    return z


def sp_pcg(m: Mesh, precond: int, co, b, x0, tol=1e-7, maxit=1000):
    x = np.array(x0, dtype=np.float64, copy=True)
    rr = ct.c_double(0.0)
    it = lib().orc_sp_pcg(ct.byref(m), precond, *_co(co), _p(np.ascontiguousarray(b)), _p(x),
                          tol, maxit, ct.byref(rr))
    return x, it, rr.value


def bicgstab2(m: Mesh, precond: int, co, b, x0, tol=1e-10, maxit=1000):
    x = np.array(x0, dtype=np.float64, copy=True)
    rr = ct.c_double(0.0)
    it = lib().orc_bicgstab2(ct.byref(m), precond, *_co(co), _p(np.ascontiguousarray(b)), _p(x),
                             tol, maxit, ct.byref(rr))
    return x, it, rr.value


def sp_finish_pressure(m: Mesh, x, p):
    p = np.array(p, dtype=np.float64, copy=True)
    lib().orc_sp_finish_pressure(ct.byref(m), _p(np.ascontiguousarray(x)), _p(p))
    return p


def sp_correct(m: Mesh, p, rho_u, rho_v, dt, u, v):
    u = np.array(u, dtype=np.float64, copy=True)
    v = np.array(v, dtype=np.float64, copy=True)
    lib().orc_sp_correct(ct.byref(m), _p(p), _p(rho_u), _p(rho_v), dt, _p(u), _p(v))
    return u, v


def bench_jacobi_pcg_solve(m: Mesh, co, b, tol: float, maxit: int, threads: int):
    """(seconds, iterations, relres) of a Jacobi-PCG solve from zero to tol on the host cores."""
    it, rr = ct.c_int(0), ct.c_double(0.0)
    sec = lib().orc_bench_jacobi_pcg_solve(ct.byref(m), *_co(co), _p(np.ascontiguousarray(b)), tol, maxit, threads,
                                           ct.byref(it), ct.byref(rr))
    return sec, it.value, rr.value


def set_threads(threads: int) -> None:
    """Threads of the reconstruction / WY cell loops (1 = serial)."""
    lib().orc_set_threads(int(threads))


def bench_jacobi_pcg(m: Mesh, co, b, iters: int, threads: int) -> float:
    return lib().orc_bench_jacobi_pcg(ct.byref(m), *_co(co), _p(b), iters, threads)


# ------------------------------------------------------- SP velocity predictor

MOM_FIELDS = ("u", "v", "u_old", "v_old", "u_nm1", "v_nm1", "rho_u", "rho_v", "mu", "mu_node", "fx", "fy")


class Momentum:
    """The SP momentum block on one mesh: fields (dict of halo arrays named as
    MOM_FIELDS), options (dict: time_scheme, has_nm1, enable_convection,
    conv_scheme, conv_lin, gx, gy, bc_u, bc_v) and the persistent work fields
    of the reference's SAVEd operator state (zero at construction)."""

    def __init__(self, m: Mesh, fields: dict, opts: dict, dt: float):
        self.m, self.dt = m, float(dt)
        self.f = {k: np.ascontiguousarray(fields[k], dtype=np.float64) for k in MOM_FIELDS}
        self.cfg = self._cfg(opts)
        self.F = MomFields(*[_p(self.f[k]) for k in MOM_FIELDS])
        shp = (m.ny + 2, m.nx + 2)
        self.tu, self.tv, self.tp = np.zeros(shp), np.zeros(shp), np.zeros(shp)
        a, b = ct.c_int(), ct.c_int()
        lib().orc_mom_sizes(ct.byref(m), ct.byref(a), ct.byref(b))
        self.u_nx, self.v_ny = a.value, b.value
        self.mono = int(opts.get("solver_type", 3)) == 2
        self.nu, self.nv = self.u_nx * m.ny, m.nx * self.v_ny
        self.n = self.nu + self.nv + (m.nx * m.ny if self.mono else 0)

    @staticmethod
    def _cfg(opts: dict) -> MomCfg:
        return MomCfg(int(opts.get("time_scheme", 1)), int(opts.get("has_nm1", 0)),
                      int(opts.get("enable_convection", 1)), int(opts.get("conv_scheme", 1)),
                      int(opts.get("conv_lin", 1)), float(opts.get("gx", 0.0)), float(opts.get("gy", 0.0)),
                      (ct.c_double * 4)(*opts.get("bc_u", (0.0,) * 4)), (ct.c_double * 4)(*opts.get("bc_v", (0.0,) * 4)),
                      int(opts.get("solver_type", 3)))

    def diag(self):
        d = np.empty(self.n)
        lib().orc_mom_diag(ct.byref(self.m), ct.byref(self.cfg), ct.byref(self.F), self.dt, _p(d))
        return d

    def rhs(self):
        r = np.zeros(self.n)
        lib().orc_mom_rhs(ct.byref(self.m), ct.byref(self.cfg), ct.byref(self.F), self.dt, _p(r))
        return r

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        if self.mono:
            lib().orc_mono_matvec(ct.byref(self.m), ct.byref(self.cfg), ct.byref(self.F), self.dt,
                                  _p(self.tu), _p(self.tv), _p(self.tp), _p(x), _p(y))
        else:
            lib().orc_mom_matvec(ct.byref(self.m), ct.byref(self.cfg), ct.byref(self.F), self.dt,
                                 _p(self.tu), _p(self.tv), _p(x), _p(y))
        return y

    def blocks(self):
        """assemble_preconditioner_blocks: dict of diag, u_a*, v_a*, schur_a*."""
        npp = self.m.nx * self.m.ny
        sizes = dict(diag=self.n, uW=self.nu, uE=self.nu, uS=self.nu, uN=self.nu, uP=self.nu, vW=self.nv,
                     vE=self.nv, vS=self.nv, vN=self.nv, vP=self.nv, sW=npp, sE=npp, sS=npp, sN=npp, sP=npp)
        arrs = {k: np.zeros(v) for k, v in sizes.items()}
        B = MomBlocks(*[_p(arrs[k]) for k, _ in MomBlocks._fields_])
        lib().orc_mom_blocks(ct.byref(self.m), ct.byref(self.cfg), ct.byref(self.F), self.dt, ct.byref(B))
        self._blocks = (arrs, B)
        return arrs

    def precond(self, precond, src):
        """apply_preconditioner (NONE, JACOBI, SSOR, BLOCK_TRIANGULAR) with
        the blocks of the fields."""
        arrs = self.blocks()
        B = self._blocks[1]
        dst = np.empty(self.n)
        lib().orc_mom_precond(ct.byref(self.m), ct.byref(self.cfg), int(precond), ct.byref(B),
                              _p(np.ascontiguousarray(src, dtype=np.float64)), _p(dst))
        return dst

    def pack(self, u, v):
        x = np.empty(self.n)
        lib().orc_mom_pack(ct.byref(self.m), _p(np.ascontiguousarray(u)), _p(np.ascontiguousarray(v)), _p(x))
        return x

    def solve(self, b, x0, lin_solver=1, precond=1, tol=1e-10, maxit=1000):
        x = np.array(x0, dtype=np.float64)
        rr = ct.c_double()
        it = lib().orc_mom_solve2(ct.byref(self.m), ct.byref(self.cfg), ct.byref(self.F), self.dt,
                                  _p(self.tu), _p(self.tv), _p(self.tp) if self.mono else None, lin_solver, precond,
                                  tol, maxit, _p(np.ascontiguousarray(b, dtype=np.float64)), _p(x), ct.byref(rr))
        return x, it, rr.value

    def mono_step(self, p, lin_solver=1, precond=4, tol=1e-10, maxit=1000):
        """The monolithic solve of sim_run: (u, v, p, iters, relres)."""
        u, v = self.f["u"].copy(), self.f["v"].copy()
        p = np.array(p, dtype=np.float64, copy=True)
        rr = ct.c_double()
        it = lib().orc_mono_step(ct.byref(self.m), ct.byref(self.cfg), ct.byref(self.F), self.dt, _p(self.tu),
                                 _p(self.tv), _p(self.tp), lin_solver, precond, tol, maxit, _p(u), _p(v), _p(p),
                                 ct.byref(rr))
        return u, v, p, it, rr.value

    def predictor(self, lin_solver=1, precond=1, tol=1e-10, maxit=1000):
        """Returns (u, v, iters, relres): the predicted velocity (halo arrays)."""
        u, v = self.f["u"].copy(), self.f["v"].copy()
        rr = ct.c_double()
        it = lib().orc_sp_predictor(ct.byref(self.m), ct.byref(self.cfg), ct.byref(self.F), self.dt,
                                    _p(self.tu), _p(self.tv), lin_solver, precond, tol, maxit,
                                    _p(u), _p(v), ct.byref(rr))
        return u, v, it, rr.value


def velocity_bc(m: Mesh, opts: dict, u, v, homogeneous=False):
    """sync_u, sync_v, apply_boundary_conditions in place (one rank)."""
    cfg = Momentum._cfg(opts)
    lib().orc_exchange_halos_scalar(ct.byref(m), _p(u))
    lib().orc_exchange_halos_scalar(ct.byref(m), _p(v))
    lib().orc_velocity_bc(ct.byref(m), ct.byref(cfg), _p(u), _p(v), int(bool(homogeneous)))


class SPRun:
    """The body of sim_run's time loop for a two-phase standard-projection run
    (src/module_simulation.f90:495-559, :2424-2666), restated over the oracle:
    n-1 / old swaps, advect_vof (WY + update_physical_properties + CSF from C),
    apply_velocity_bc, the momentum block, the Poisson block (CG + SGS, 1e-7,
    1000 iterations, mean removal), the velocity correction, sync + BC.  The
    initial properties are the reference's reset values, then one
    update_physical_properties.  opts: the momentum options (see Momentum)
    with conv_lin etc.; the momentum operator's persistent work fields carry
    over from step to step, as the reference's do."""

    def __init__(self, m: Mesh, C, u, v, rho1, rho2, mu1, mu2, sigma, opts: dict, method=PLIC,
                 lin_solver=3, precond=1, tol=1e-12, maxit=500):
        self.m, self.opts, self.method = m, dict(opts), method
        self.rho1, self.rho2, self.mu1, self.mu2, self.sigma = rho1, rho2, mu1, mu2, sigma
        self.solver = (lin_solver, precond, tol, maxit)
        shp = (m.ny + 2, m.nx + 2)
        self.C = np.array(C, dtype=np.float64, copy=True)
        self.u = np.array(u, dtype=np.float64, copy=True)
        self.v = np.array(v, dtype=np.float64, copy=True)
        self.p = np.zeros(shp)
        self.rho, self.rho_u, self.rho_v = np.ones(shp), np.ones(shp), np.ones(shp)
        self.mu, self.mu_node = np.full(shp, 1.0e-3), np.full(shp, 1.0e-3)
        self.fx, self.fy, self.curv = np.zeros(shp), np.zeros(shp), np.zeros(shp)
        self.u_old, self.v_old = np.zeros(shp), np.zeros(shp)
        self.u_nm1, self.v_nm1 = np.zeros(shp), np.zeros(shp)
        self.tu, self.tv = np.zeros(shp), np.zeros(shp)
        self.step_count = 0
        self.nstep = 0
        self._props()

    def _props(self):
        lib().orc_update_properties(ct.byref(self.m), _p(self.C), self.rho1, self.rho2, self.mu1, self.mu2,
                                    _p(self.rho), _p(self.rho_u), _p(self.rho_v), _p(self.mu), _p(self.mu_node))

    def step(self, dt):
        m = self.m
        self.nstep += 1
        if self.nstep > 1:
            self.u_nm1, self.v_nm1 = self.u_old.copy(), self.v_old.copy()
        else:
            self.u_nm1, self.v_nm1 = self.u.copy(), self.v.copy()
        has_nm1 = self.nstep > 1
        self.u_old, self.v_old = self.u.copy(), self.v.copy()
        # advect_vof: transport, properties, CSF (normals and gradient from C)
        self.C, self.step_count = wy_step(m, self.method, self.C, self.u, self.v, dt, self.step_count)
        self._props()
        lib().orc_curvature_csf(ct.byref(m), _p(self.C), _p(self.C), int(self.sigma > 0.0), self.sigma,
                                _p(self.curv), _p(self.fx), _p(self.fy))
        velocity_bc(m, self.opts, self.u, self.v, False)
        # the momentum block
        opts = dict(self.opts, has_nm1=int(has_nm1))
        fields = dict(u=self.u, v=self.v, u_old=self.u_old, v_old=self.v_old, u_nm1=self.u_nm1, v_nm1=self.v_nm1,
                      rho_u=self.rho_u, rho_v=self.rho_v, mu=self.mu, mu_node=self.mu_node, fx=self.fx, fy=self.fy)
        M = Momentum(m, fields, opts, dt)
        M.tu, M.tv = self.tu, self.tv
        self.u, self.v, it_u, _ = M.predictor(*self.solver)
        # the Poisson block and the correction
        co = poisson_assemble(m, self.u, self.v, self.rho_u, self.rho_v, dt)
        x, it_p, _ = sp_pcg(m, 2, co, co["b"], self.p[1:m.ny + 1, 1:m.nx + 1].copy(), tol=1e-7, maxit=1000)
        self.p = sp_finish_pressure(m, x, self.p)
        self.u, self.v = sp_correct(m, self.p, self.rho_u, self.rho_v, dt, self.u, self.v)
        velocity_bc(m, self.opts, self.u, self.v, False)
        return it_u + it_p


def history(m: Mesh, C, u, v, p, rho, T=None):
    """vtk_write_diagnostics' eight quantities (serial sums)."""
    out = np.zeros(8)
    arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (C, u, v, p, rho)]
    Tp = _p(np.ascontiguousarray(T, dtype=np.float64)) if T is not None else None
    lib().orc_history(ct.byref(m), *[_p(a) for a in arr], Tp, _p(out))
    return out
