"""Host-side mirror of the reference's hot-path interface over libcovof_hip."""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass

import torch

from . import _lib as L


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


@dataclass
class SolveResult:
    """Outcome of a Krylov solve: iterations (inner ones summed in mixed
    precision), final relative residual, status code (include/covof_hip.h)."""
    iters: int
    relres: float
    status: int

    @property
    def converged(self) -> bool:
        return self.status == L.OK


class Context:
    """One mesh + its device buffers (``covof_ctx_create``).

    ``bc`` = (left, right, bottom, top) reference boundary types
    (src/module_constants.f90:25-31).  Calls run on torch's current CUDA
    stream of the calling thread.
    """

    def __init__(self, nx: int, ny: int, lx: float = 1.0, ly: float = 1.0,
                 bc=(L.BC_PERIODIC,) * 4, device: int | str | torch.device = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("covof_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device("cuda", torch.device(device).index or 0) if not isinstance(device, int) \
            else torch.device("cuda", device)
        self.nx, self.ny, self.lx, self.ly = int(nx), int(ny), float(lx), float(ly)
        self.bc = tuple(int(b) for b in bc)
        self._lib = L.lib()
        self._h = ct.c_void_p()
        arr = (ct.c_int * 4)(*self.bc)
        with torch.cuda.device(self.device):
            L.check(self._lib.covof_ctx_create(self.nx, self.ny, self.lx, self.ly, arr, ct.byref(self._h)),
                    "covof_ctx_create")
        self.step_count = 0  # the reference's vof_step_count (src/module_interface_vof.f90:27)

    # ------------------------------------------------------------ plumbing
    def close(self):
        if self._h:
            self._lib.covof_ctx_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        L.check(self._lib.covof_ctx_set_stream(self._h, ct.c_void_p(s)), "covof_ctx_set_stream")

    def _field(self, t: torch.Tensor, name: str, shape=None):
        shape = shape or (self.ny + 2, self.nx + 2)
        if not (isinstance(t, torch.Tensor) and t.device.type == "cuda" and t.dtype == torch.float64
                and tuple(t.shape) == tuple(shape) and t.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous float64 CUDA tensor of shape {shape}")
        return t

    def _work(self, t, name):
        return self._field(t, name, (self.ny, self.nx))

    def new_field(self, fill=0.0):
        return torch.full((self.ny + 2, self.nx + 2), fill, dtype=torch.float64, device=self.device)

    def new_work(self, fill=0.0):
        return torch.full((self.ny, self.nx), fill, dtype=torch.float64, device=self.device)

    def profile(self, period: int):
        """HIP-event timing of the hot kernels (covof_ctx_profile): every VOF
        launch, Krylov iterations sampled every ``period``-th; 0 turns it off."""
        L.check(self._lib.covof_ctx_profile(self._h, int(period)), "covof_ctx_profile")

    def profile_read(self, tag: int):
        """(total ms, launches) of one kernel tag since profiling was turned on."""
        ms, n = ct.c_double(0.0), ct.c_long(0)
        L.check(self._lib.covof_ctx_profile_read(self._h, int(tag), ct.byref(ms), ct.byref(n)),
                "covof_ctx_profile_read")
        return ms.value, n.value

    def set_reduction_order(self, order: int):
        """Summation order of the FP64 Krylov inner products and the pressure
        mean (covof_ctx_set_reduction_order): REDUCTION_TREE (default) or
        REDUCTION_REFERENCE, the reference's serial column-major sums
        (src/module_simulation.f90:2760-2774), which reproduce its iterates
        bit for bit."""
        L.check(self._lib.covof_ctx_set_reduction_order(self._h, int(order)), "covof_ctx_set_reduction_order")

    def synchronize(self):
        L.check(self._lib.covof_ctx_synchronize(self._h), "covof_ctx_synchronize")

    # ------------------------------------------------------------ geometry
    def geometry_batch(self, xmin, xmax, ymin, ymax, nx, ny, alpha, c):
        """cut_box_area and compute_alpha over 1-D float64 CUDA tensors
        (src/module_interface_reconstruction.f90:33-147)."""
        ins = [torch.as_tensor(a, dtype=torch.float64, device=self.device).contiguous()
               for a in (xmin, xmax, ymin, ymax, nx, ny, alpha, c)]
        n = ins[0].numel()
        area = torch.empty(n, dtype=torch.float64, device=self.device)
        aout = torch.empty_like(area)
        self._stream()
        L.check(self._lib.covof_geom_batch_d(self._h, n, *[_ptr(a) for a in ins], _ptr(area), _ptr(aout)),
                "covof_geom_batch_d")
        return area, aout

    def reconstruct_interface(self, method: str, C: torch.Tensor):
        """reconstruct_interface(method_name, mesh, C, segments)
        (src/module_interface_reconstruction.f90:580-683) -> (nx, ny, alpha) fields."""
        self._field(C, "C")
        out = [self.new_field() for _ in range(3)]
        self._stream()
        L.check(self._lib.covof_reconstruct_interface_d(self._h, method.encode(), _ptr(C),
                                                        *[_ptr(o) for o in out]),
                "covof_reconstruct_interface_d")
        return tuple(out)

    # ---------------------------------------------------------------- VOF
    def advect_vof(self, C, u, v, dt, advection_method="WEYMOUTH_YUE", reconstruction_method="PLIC",
                   nsteps=1, two_phase=False, rho1=1.0, rho2=1.0, mu1=1.0e-3, mu2=1.0e-3, props=None,
                   step_count=None):
        """advect_vof(fields, mesh, cfg, dt) (src/module_interface_vof.f90:792-825), ``nsteps``
        times, C in place.  ``props`` = dict of rho, rho_u, rho_v, mu, mu_node fields
        (written when ``two_phase``).  Uses/advances ``self.step_count`` unless given."""
        for t, n in ((C, "C"), (u, "u"), (v, "v")):
            self._field(t, n)
        pp = [None] * 5
        if two_phase:
            if props is None:
                props = {k: self.new_field(1.0) for k in ("rho", "rho_u", "rho_v", "mu", "mu_node")}
            pp = [self._field(props[k], k) for k in ("rho", "rho_u", "rho_v", "mu", "mu_node")]
        sc = ct.c_int(self.step_count if step_count is None else int(step_count))
        self._stream()
        L.check(self._lib.covof_vof_advect_d(self._h, _ptr(C), _ptr(u), _ptr(v), float(dt),
                                             advection_method.encode(), reconstruction_method.encode(),
                                             ct.byref(sc), int(nsteps), int(bool(two_phase)), float(rho1),
                                             float(rho2), float(mu1), float(mu2), *[_ptr(p) for p in pp]),
                "covof_vof_advect_d")
        if step_count is None:
            self.step_count = sc.value
        return props if two_phase else None

    def update_physical_properties(self, C, rho1, rho2, mu1, mu2, props=None, C_s=None):
        """update_physical_properties, two-phase incompressible branch
        (src/module_interface_vof.f90:479-616); fixes C's physical corners in
        place.  ``C_s``: the SVOF-smoothed field (cfg%enable_svof_smoothing):
        rho, mu and mu_node then come from it (:540-545, :596-600)."""
        self._field(C, "C")
        if C_s is not None:
            self._field(C_s, "C_s")
        if props is None:
            props = {k: self.new_field(1.0) for k in ("rho", "rho_u", "rho_v", "mu", "mu_node")}
        pp = [self._field(props[k], k) for k in ("rho", "rho_u", "rho_v", "mu", "mu_node")]
        self._stream()
        L.check(self._lib.covof_update_properties_cs_d(self._h, _ptr(C), _ptr(C_s), float(rho1), float(rho2),
                                                       float(mu1), float(mu2), *[_ptr(p) for p in pp]),
                "covof_update_properties_cs_d")
        return props

    def set_solver_option(self, option: int, value: float):
        """covof_ctx_set_solver_option: one of the OPT_* tunings of the
        Poisson solves (drops the captured cycles and coarse operators)."""
        L.check(self._lib.covof_ctx_set_solver_option(self._h, int(option), float(value)),
                "covof_ctx_set_solver_option")

    def reserve(self, parts: int = L.RESERVE_ALL):
        """covof_ctx_reserve: make the lazily made work arrays now (an OR of
        RESERVE_* parts), so the calls of a time step allocate nothing."""
        L.check(self._lib.covof_ctx_reserve(self._h, int(parts)), "covof_ctx_reserve")

    def mixed_solve_info(self):
        """(FP32 iterations, FP64 continuation iterations) of the last
        mixed-precision CG solve."""
        a, b = ct.c_int(0), ct.c_int(0)
        L.check(self._lib.covof_ctx_mixed_solve_info(self._h, ct.byref(a), ct.byref(b)), "covof_ctx_mixed_solve_info")
        return a.value, b.value

    def set_obstacle(self, obs_type=None, x0=0.0, y0=0.0, x1=0.0, y1=0.0, xc=0.0, yc=0.0, radius=0.0,
                     perm=0.0):
        """The solid obstacle of config_t (obs_type 1 box [x0,x1]x[y0,y1], 2
        cylinder (xc, yc, radius); obs_perm) the Poisson assembly and the
        velocity correction see (src/module_simulation.f90:2491-2528,
        :2596-2632); obs_type None removes it."""
        if obs_type is None:
            L.check(self._lib.covof_ctx_set_obstacle(self._h, None), "covof_ctx_set_obstacle")
            return
        o = L.Obstacle(1, int(obs_type), x0, y0, x1, y1, xc, yc, radius, perm)
        L.check(self._lib.covof_ctx_set_obstacle(self._h, ct.byref(o)), "covof_ctx_set_obstacle")

    def set_vof_top_profile(self, profile="UNIFORM", jet_center=0.5, jet_width=0.08):
        """The top side's VOF profile of apply_vof_bc (bc_top%profile,
        jet_center, jet_width; src/module_mpi_comm.f90:664-676)."""
        L.check(self._lib.covof_ctx_set_vof_top_profile(self._h, profile.encode(), float(jet_center),
                                                        float(jet_width)), "covof_ctx_set_vof_top_profile")

    # ------------------------------------------------ after the transport
    def compute_curvature_and_csf(self, C, sigma, G=None, enable=True, out=None, method="normals"):
        """compute_curvature_and_csf (src/module_interface_vof.f90:627-789):
        returns {"curvature", "fx_sigma", "fy_sigma"} (halo fields); normals
        from G (C_tilde / C_s when a smoothing is on; default C).  ``out``
        supplies the three arrays (the force arrays keep their entries outside
        the face ranges); by default they start at zero.  ``method``:
        "normals" (the reference's -div(n)) or "height_function" (3 x 7
        column heights of C where they bracket the interface, the reference's
        value elsewhere; no reference counterpart)."""
        codes = {"normals": L.CURVATURE_NORMALS, "height_function": L.CURVATURE_HEIGHT_FUNCTION}
        if method not in codes:
            raise ValueError(f"method must be one of {sorted(codes)}")
        self._field(C, "C")
        if G is not None:
            self._field(G, "G")
        if out is None:
            out = {k: self.new_field(0.0) for k in ("curvature", "fx_sigma", "fy_sigma")}
        arrs = [self._field(out[k], k) for k in ("curvature", "fx_sigma", "fy_sigma")]
        self._stream()
        L.check(self._lib.covof_curvature_csf_method_d(self._h, _ptr(G), _ptr(C), 1 if enable else 0, float(sigma),
                                                       codes[method], *[_ptr(a) for a in arrs]),
                "covof_curvature_csf_method_d")
        return out

    def smooth_diffusion(self, C, steps=3, C_tilde=None):
        """vof_smooth_diffusion (src/module_interface_vof.f90:1416-1461): C_tilde."""
        self._field(C, "C")
        C_tilde = self.new_field(0.0) if C_tilde is None else self._field(C_tilde, "C_tilde")
        self._stream()
        L.check(self._lib.covof_smooth_diffusion_d(self._h, _ptr(C), int(steps), _ptr(C_tilde)),
                "covof_smooth_diffusion_d")
        return C_tilde

    def smooth_svof_helmholtz(self, C, filter_width=1.0, C_s=None):
        """smooth_svof_helmholtz (src/module_interface_vof.f90:387-476): C_s."""
        self._field(C, "C")
        C_s = self.new_field(0.0) if C_s is None else self._field(C_s, "C_s")
        self._stream()
        L.check(self._lib.covof_svof_smooth_d(self._h, _ptr(C), float(filter_width), _ptr(C_s)),
                "covof_svof_smooth_d")
        return C_s

    # ------------------------------------------------------------ Poisson
    def poisson_assemble(self, u, v, rho_u, rho_v, dt, b=None):
        """SP Poisson assembly (src/module_simulation.f90:2477-2533); returns b = -div(u*)."""
        for t, n in ((u, "u"), (v, "v"), (rho_u, "rho_u"), (rho_v, "rho_v")):
            self._field(t, n)
        b = self.new_work() if b is None else self._work(b, "b")
        self._stream()
        L.check(self._lib.covof_poisson_assemble_d(self._h, _ptr(u), _ptr(v), _ptr(rho_u), _ptr(rho_v),
                                                   float(dt), _ptr(b)), "covof_poisson_assemble_d")
        return b

    def poisson_matvec(self, x, y=None):
        """poisson_matvec (src/module_simulation.f90:2668-2707)."""
        self._work(x, "x")
        y = self.new_work() if y is None else self._work(y, "y")
        self._stream()
        L.check(self._lib.covof_poisson_matvec_d(self._h, _ptr(x), _ptr(y)), "covof_poisson_matvec_d")
        return y

    def poisson_precond(self, precond, r, z=None):
        """NONE / JACOBI / SSOR (= poisson_ssor, src/module_simulation.f90:2709-2758)."""
        self._work(r, "r")
        z = self.new_work() if z is None else self._work(z, "z")
        self._stream()
        L.check(self._lib.covof_poisson_precond_d(self._h, int(precond), _ptr(r), _ptr(z)),
                "covof_poisson_precond_d")
        return z

    def poisson_solve(self, b, x, linear_solver=L.LINEAR_CG, precond=L.PRECOND_JACOBI, tol=1e-10,
                      max_iter=100000, precision=L.PRECISION_FP64) -> SolveResult:
        """Krylov solve of A x = b on the assembled operator (x in/out)."""
        self._work(b, "b")
        self._work(x, "x")
        it, rr = ct.c_int(0), ct.c_double(0.0)
        self._stream()
        st = self._lib.covof_poisson_solve_d(self._h, _ptr(b), _ptr(x), int(linear_solver), int(precond),
                                             float(tol), int(max_iter), int(precision), ct.byref(it),
                                             ct.byref(rr))
        L.check(st, "covof_poisson_solve_d", allow=(L.OK, L.NOT_CONVERGED, L.BREAKDOWN, L.NONFINITE))
        return SolveResult(it.value, rr.value, st)

    def solve_standard_projection_poisson(self, u, v, rho_u, rho_v, p, dt, linear_solver=L.LINEAR_CG,
                                          precond=L.PRECOND_SSOR, tol=1e-7, max_iter=1000,
                                          precision=L.PRECISION_FP64) -> SolveResult:
        """The Poisson block of sim_solve_standard_projection
        (src/module_simulation.f90:2477-2588), p in/out; defaults are the
        reference's hard-coded CG + SSOR, tol 1e-7, 1000 iterations."""
        for t, n in ((u, "u"), (v, "v"), (rho_u, "rho_u"), (rho_v, "rho_v"), (p, "p")):
            self._field(t, n)
        it, rr = ct.c_int(0), ct.c_double(0.0)
        self._stream()
        st = self._lib.covof_sp_poisson_d(self._h, _ptr(u), _ptr(v), _ptr(rho_u), _ptr(rho_v), _ptr(p),
                                          float(dt), int(linear_solver), int(precond), float(tol),
                                          int(max_iter), int(precision), ct.byref(it), ct.byref(rr))
        L.check(st, "covof_sp_poisson_d", allow=(L.OK, L.NOT_CONVERGED, L.BREAKDOWN, L.NONFINITE))
        return SolveResult(it.value, rr.value, st)

    def correct_velocity(self, p, rho_u, rho_v, dt, u, v):
        """Projection u -= (dt/rho_u) dp/dx, v -= (dt/rho_v) dp/dy
        (src/module_simulation.f90:2590-2654), u and v in place."""
        for t, n in ((p, "p"), (rho_u, "rho_u"), (rho_v, "rho_v"), (u, "u"), (v, "v")):
            self._field(t, n)
        self._stream()
        L.check(self._lib.covof_sp_correct_d(self._h, _ptr(p), _ptr(rho_u), _ptr(rho_v), float(dt), _ptr(u),
                                             _ptr(v)), "covof_sp_correct_d")

    def poisson_adjoint(self, rho, dt, x, w, linear_solver=L.LINEAR_CG, precond=L.PRECOND_JACOBI,
                        tol=1e-12, max_iter=200000, precision=L.PRECISION_FP64):
        """dJ/db and dJ/d(log rho) for w = dJ/dx at the solution x of the
        assembled operator (no reference counterpart)."""
        self._field(rho, "rho")
        self._work(x, "x")
        self._work(w, "w")
        g_b, g_lr = self.new_work(), self.new_work()
        it, rr = ct.c_int(0), ct.c_double(0.0)
        self._stream()
        st = self._lib.covof_poisson_adjoint_d(self._h, _ptr(rho), float(dt), _ptr(x), _ptr(w), _ptr(g_b),
                                               _ptr(g_lr), int(linear_solver), int(precond), float(tol),
                                               int(max_iter), int(precision), ct.byref(it), ct.byref(rr))
        L.check(st, "covof_poisson_adjoint_d", allow=(L.OK, L.NOT_CONVERGED, L.BREAKDOWN, L.NONFINITE))
        return g_b, g_lr, SolveResult(it.value, rr.value, st)

    def sp_pressure(self, x, p):
        """p(1:nx,1:ny) = x - mean(x) with the periodic halo wrap, the end of
        the SP Poisson block (src/module_simulation.f90:2577-2588)."""
        self._work(x, "x")
        self._field(p, "p")
        self._stream()
        L.check(self._lib.covof_sp_pressure_d(self._h, _ptr(x), _ptr(p)), "covof_sp_pressure_d")

    def projection_adjoint(self, rho, rho_u, rho_v, dt, x, w_u, w_v, w_p=None, precond=L.PRECOND_JACOBI,
                           tol=1e-12, max_iter=200000, precision=L.PRECISION_FP64):
        """dJ/du*, dJ/dv* (halo fields) and dJ/d(log rho) (work array) of the
        whole projection (divergence, solve, mean removal, correction) for the
        objective's gradients w_u, w_v, w_p at the forward's solution x; the
        operator must be the forward's (no reference counterpart)."""
        for t, n in ((rho, "rho"), (rho_u, "rho_u"), (rho_v, "rho_v"), (w_u, "w_u"), (w_v, "w_v")):
            self._field(t, n)
        if w_p is not None:
            self._field(w_p, "w_p")
        self._work(x, "x")
        du, dv, dlr = self.new_field(), self.new_field(), self.new_work()
        it, rr = ct.c_int(0), ct.c_double(0.0)
        self._stream()
        st = self._lib.covof_projection_adjoint_d(self._h, _ptr(rho), _ptr(rho_u), _ptr(rho_v), float(dt), _ptr(x),
                                                  _ptr(w_u), _ptr(w_v), _ptr(w_p), _ptr(du), _ptr(dv), _ptr(dlr),
                                                  int(precond), float(tol), int(max_iter), int(precision),
                                                  ct.byref(it), ct.byref(rr))
        L.check(st, "covof_projection_adjoint_d", allow=(L.OK, L.NOT_CONVERGED, L.BREAKDOWN, L.NONFINITE))
        return du, dv, dlr, SolveResult(it.value, rr.value, st)

    # ------------------------------- pieces of the multi-step kinetic-energy adjoint
    def vof_advect_adjoint(self, C, u, v, dt, dJ_dCnew, step_count, reconstruction_method="PLIC", two_phase=True,
                           fold_input_halo=False, dJ_dC=None, dJ_du=None, dJ_dv=None):
        """Reverse mode of one Weymouth-Yue step of advect_vof from C (the
        step's input; pass order from ``step_count``): returns (dJ/dC^n,
        dJ/du, dJ/dv) for dJ_dCnew = dJ/dC^{n+1}; dJ/du and dJ/dv are added
        to ``dJ_du``, ``dJ_dv`` when given (new zero fields otherwise).
        ``fold_input_halo``: C's halo is the copy of its interior, and its
        entries are folded onto the cells they copy (no reference
        counterpart)."""
        for t, n in ((C, "C"), (u, "u"), (v, "v"), (dJ_dCnew, "dJ_dCnew")):
            self._field(t, n)
        dJ_dC = self.new_field() if dJ_dC is None else self._field(dJ_dC, "dJ_dC")
        dJ_du = self.new_field() if dJ_du is None else self._field(dJ_du, "dJ_du")
        dJ_dv = self.new_field() if dJ_dv is None else self._field(dJ_dv, "dJ_dv")
        self._stream()
        L.check(self._lib.covof_vof_advect_adjoint_d(self._h, _ptr(C), _ptr(u), _ptr(v), float(dt),
                                                     reconstruction_method.encode(), int(step_count),
                                                     int(bool(two_phase)), int(bool(fold_input_halo)),
                                                     _ptr(dJ_dCnew), _ptr(dJ_dC), _ptr(dJ_du), _ptr(dJ_dv)),
                "covof_vof_advect_adjoint_d")
        return dJ_dC, dJ_du, dJ_dv

    def fill_vof_halo(self, C):
        """C's halo from its interior as two-phase transport leaves it
        (zero-gradient copies or wrap, corners from the diagonal cell), in place."""
        self._field(C, "C")
        self._stream()
        L.check(self._lib.covof_vof_fill_halo_d(self._h, _ptr(C)), "covof_vof_fill_halo_d")
        return C

    def kinetic_energy(self, C, u, v, rho1, rho2, grad=False):
        """The history row's kinetic energy (src/module_vtk_writer.f90:399-531)
        with rho from C; with ``grad`` also (dKE/dC, dKE/du, dKE/dv)."""
        for t, n in ((C, "C"), (u, "u"), (v, "v")):
            self._field(t, n)
        ke = ct.c_double(0.0)
        g = [self.new_field() for _ in range(3)] if grad else [None] * 3
        self._stream()
        L.check(self._lib.covof_kinetic_energy_d(self._h, _ptr(C), _ptr(u), _ptr(v), float(rho1), float(rho2),
                                                 ct.byref(ke), *[_ptr(a) for a in g]), "covof_kinetic_energy_d")
        return (ke.value, *g) if grad else ke.value

    def properties_adjoint(self, C, rho1, rho2, dJ_dlogrho, dJ_dC):
        """dJ_dC += dJ/d(log rho) * drho/dC of update_physical_properties's
        density (no reference counterpart)."""
        self._field(C, "C")
        self._work(dJ_dlogrho, "dJ_dlogrho")
        self._field(dJ_dC, "dJ_dC")
        self._stream()
        L.check(self._lib.covof_properties_adjoint_d(self._h, _ptr(C), float(rho1), float(rho2), _ptr(dJ_dlogrho),
                                                     _ptr(dJ_dC)), "covof_properties_adjoint_d")
        return dJ_dC

    def inertial_predictor(self, first, dt, gx, gy, u, v, u_nm1=None, v_nm1=None, u_star=None, v_star=None):
        """u*, v* of the momentum block with convection and viscosity off:
        BDF2 inertia and gravity, explicit Euler when ``first``."""
        for t, n in ((u, "u"), (v, "v")):
            self._field(t, n)
        if not first:
            self._field(u_nm1, "u_nm1")
            self._field(v_nm1, "v_nm1")
        u_star = self.new_field() if u_star is None else self._field(u_star, "u_star")
        v_star = self.new_field() if v_star is None else self._field(v_star, "v_star")
        self._stream()
        L.check(self._lib.covof_inertial_predictor_d(self._h, int(bool(first)), float(dt), float(gx), float(gy),
                                                     _ptr(u), _ptr(v), _ptr(u_nm1), _ptr(v_nm1), _ptr(u_star),
                                                     _ptr(v_star)), "covof_inertial_predictor_d")
        return u_star, v_star

    def inertial_predictor_adjoint(self, first, w_u, w_v, dJ_du, dJ_dv, dJ_du_nm1=None, dJ_dv_nm1=None):
        """Transpose of inertial_predictor, added to the gradients."""
        for t, n in ((w_u, "w_u"), (w_v, "w_v"), (dJ_du, "dJ_du"), (dJ_dv, "dJ_dv")):
            self._field(t, n)
        if not first:
            self._field(dJ_du_nm1, "dJ_du_nm1")
            self._field(dJ_dv_nm1, "dJ_dv_nm1")
        self._stream()
        L.check(self._lib.covof_inertial_predictor_adjoint_d(self._h, int(bool(first)), _ptr(w_u), _ptr(w_v),
                                                             _ptr(dJ_du), _ptr(dJ_dv), _ptr(dJ_du_nm1),
                                                             _ptr(dJ_dv_nm1)), "covof_inertial_predictor_adjoint_d")

    # ------------------------------------------------- SP velocity predictor
    MOMENTUM_FIELDS = ("u", "v", "u_old", "v_old", "u_nm1", "v_nm1", "rho_u", "rho_v", "mu", "mu_node",
                       "fx_sigma", "fy_sigma")

    def momentum_size(self, solver_type=None):
        """(n, u_nx, v_ny) of the momentum unknowns (ns_init,
        src/module_navier_stokes.f90:136-226): u, v, and p when monolithic
        (solver_type; default: that of the bound fields)."""
        if solver_type is None:
            solver_type = getattr(self, "_mom_solver", L.SOLVER_STANDARD_PROJECTION)
        n, a, b = ct.c_int(), ct.c_int(), ct.c_int()
        L.check(self._lib.covof_momentum_size(self._h, int(solver_type), ct.byref(n), ct.byref(a), ct.byref(b)),
                "covof_momentum_size")
        return n.value, a.value, b.value

    @staticmethod
    def momentum_cfg(time_scheme=L.TIME_SCHEME_EULER, has_nm1=False, enable_convection=True,
                     convective_scheme=L.CONV_UPWIND, convective_linearization=L.CONV_LIN_PICARD, gx=0.0, gy=0.0,
                     bc_u=(0.0,) * 4, bc_v=(0.0,) * 4, solver_type=L.SOLVER_STANDARD_PROJECTION) -> L.MomentumCfg:
        """The config_t fields the momentum block reads (reference defaults:
        EULER, convection on, UPWIND, PICARD; solver_type STANDARD_PROJECTION
        or MONOLITHIC)."""
        return L.MomentumCfg(int(time_scheme), int(bool(has_nm1)), int(bool(enable_convection)),
                             int(convective_scheme), int(convective_linearization), float(gx), float(gy),
                             (ct.c_double * 4)(*bc_u), (ct.c_double * 4)(*bc_v), int(solver_type))

    def _mom_ptrs(self, fields: dict):
        out = []
        for k in self.MOMENTUM_FIELDS:
            t = fields.get(k)
            if t is None and k in ("u_nm1", "v_nm1"):
                out.append(None)
                continue
            out.append(_ptr(self._field(t, k)))
        return out

    def momentum_set_fields(self, cfg: L.MomentumCfg, fields: dict, dt: float):
        """Binds the fields_old of the momentum block (tensors named as
        MOMENTUM_FIELDS; u_nm1 / v_nm1 may be absent without an n-1 state).
        The tensors must stay alive while the block is used."""
        self._mom_keep = (cfg, dict(fields))
        self._stream()
        L.check(self._lib.covof_momentum_set_fields_d(self._h, ct.byref(cfg), *self._mom_ptrs(fields), float(dt)),
                "covof_momentum_set_fields_d")
        self._mom_solver = cfg.solver_type or L.SOLVER_STANDARD_PROJECTION

    def _mom_vec(self, t, name):
        n = self.momentum_size()[0]
        if not (isinstance(t, torch.Tensor) and t.device.type == "cuda" and t.dtype == torch.float64
                and t.numel() == n and t.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous float64 CUDA tensor of {n} entries")
        return t

    def new_momentum_vector(self, fill=0.0):
        return torch.full((self.momentum_size()[0],), float(fill), dtype=torch.float64, device=self.device)

    def assemble_preconditioner_blocks(self, diag=None):
        """diag_u, diag_v (and diag_p when monolithic) of
        assemble_preconditioner_blocks (src/module_navier_stokes.f90:1268-1595)
        of the bound fields."""
        diag = self.new_momentum_vector() if diag is None else self._mom_vec(diag, "diag")
        self._stream()
        L.check(self._lib.covof_momentum_diag_d(self._h, _ptr(diag)), "covof_momentum_diag_d")
        return diag

    def assemble_rhs(self, rhs=None):
        """ns_assemble_rhs (src/module_navier_stokes.f90:785-1101) of the bound fields."""
        rhs = self.new_momentum_vector() if rhs is None else self._mom_vec(rhs, "rhs")
        self._stream()
        L.check(self._lib.covof_momentum_rhs_d(self._h, _ptr(rhs)), "covof_momentum_rhs_d")
        return rhs

    def apply_system_matrix(self, x, y=None):
        """y = A x, ns_apply_system_matrix (src/module_navier_stokes.f90:380-780)."""
        self._mom_vec(x, "x")
        y = self.new_momentum_vector() if y is None else self._mom_vec(y, "y")
        self._stream()
        L.check(self._lib.covof_momentum_matvec_d(self._h, _ptr(x), _ptr(y)), "covof_momentum_matvec_d")
        return y

    def solve_linear_system(self, b, x, linear_solver=L.LINEAR_BICGSTAB2, precond=L.PRECOND_JACOBI, tol=1e-5,
                            max_iter=300) -> SolveResult:
        """solve_linear_system (src/module_linear_solvers.f90:1167-1208) on the
        momentum operator, x in/out; defaults are config_t's (tol 1e-5, 300)."""
        self._mom_vec(b, "b")
        self._mom_vec(x, "x")
        it, rr = ct.c_int(0), ct.c_double(0.0)
        self._stream()
        st = self._lib.covof_momentum_solve_d(self._h, _ptr(b), _ptr(x), int(linear_solver), int(precond),
                                              float(tol), int(max_iter), ct.byref(it), ct.byref(rr))
        L.check(st, "covof_momentum_solve_d", allow=(L.OK, L.NOT_CONVERGED, L.BREAKDOWN, L.NONFINITE))
        return SolveResult(it.value, rr.value, st)

    def apply_preconditioner(self, r, z=None, precond=L.PRECOND_JACOBI):
        """z = M^-1 r, apply_preconditioner (src/module_linear_solvers.f90:
        48-316) with the blocks of the bound fields: NONE, JACOBI, SSOR or
        BLOCK_TRIANGULAR."""
        self._mom_vec(r, "r")
        z = self.new_momentum_vector() if z is None else self._mom_vec(z, "z")
        self._stream()
        L.check(self._lib.covof_momentum_precond_d(self._h, int(precond), _ptr(r), _ptr(z)),
                "covof_momentum_precond_d")
        return z

    def apply_velocity_bc(self, cfg: L.MomentumCfg, u, v, homogeneous=False):
        """sync_u, sync_v, apply_boundary_conditions (src/module_mpi_comm.f90:
        733-775, :290-537) on the halo fields u, v in place."""
        self._field(u, "u")
        self._field(v, "v")
        self._stream()
        L.check(self._lib.covof_velocity_bc_d(self._h, ct.byref(cfg), _ptr(u), _ptr(v), int(bool(homogeneous))),
                "covof_velocity_bc_d")

    def solve_standard_projection_predictor(self, cfg: L.MomentumCfg, fields: dict, dt: float,
                                            linear_solver=L.LINEAR_BICGSTAB2, precond=L.PRECOND_JACOBI,
                                            tol=1e-5, max_iter=300) -> SolveResult:
        """The momentum block of sim_solve_standard_projection
        (src/module_simulation.f90:2454-2469): fields["u"], fields["v"] are
        replaced in place by the predicted velocity."""
        it, rr = ct.c_int(0), ct.c_double(0.0)
        self._stream()
        st = self._lib.covof_sp_predictor_d(self._h, ct.byref(cfg), *self._mom_ptrs(fields), float(dt),
                                            int(linear_solver), int(precond), float(tol), int(max_iter),
                                            ct.byref(it), ct.byref(rr))
        L.check(st, "covof_sp_predictor_d", allow=(L.OK, L.NOT_CONVERGED, L.BREAKDOWN, L.NONFINITE))
        self._mom_solver = L.SOLVER_STANDARD_PROJECTION
        return SolveResult(it.value, rr.value, st)

    def solve_monolithic_step(self, cfg: L.MomentumCfg, fields: dict, p, dt: float,
                              linear_solver=L.LINEAR_BICGSTAB2, precond=L.PRECOND_BLOCK_TRIANGULAR, tol=1e-5,
                              max_iter=300) -> SolveResult:
        """The monolithic branch of the time step (src/module_simulation.f90:
        553-583, one Uzawa pass): fields["u"], fields["v"] and p (halo
        tensors) are replaced in place by the coupled solution.  cfg must
        have solver_type MONOLITHIC."""
        self._field(p, "p")
        it, rr = ct.c_int(0), ct.c_double(0.0)
        self._stream()
        ptrs = self._mom_ptrs(fields)
        st = self._lib.covof_monolithic_step_d(self._h, ct.byref(cfg), ptrs[0], ptrs[1], _ptr(p), *ptrs[2:],
                                               float(dt), int(linear_solver), int(precond), float(tol),
                                               int(max_iter), ct.byref(it), ct.byref(rr))
        L.check(st, "covof_monolithic_step_d", allow=(L.OK, L.NOT_CONVERGED, L.BREAKDOWN, L.NONFINITE))
        self._mom_solver = L.SOLVER_MONOLITHIC
        return SolveResult(it.value, rr.value, st)

    # ------------------------------------------------------------- output
    def history_row(self, C, u, v, p, rho, T=None):
        """The quantities of one _history.csv row (vtk_write_diagnostics,
        src/module_vtk_writer.f90:399-531): [center_of_mass_y, rise_velocity,
        bubble_volume, kinetic_energy, max_u, max_div, avg_rho_gas,
        avg_p_gas].  T None: 300 everywhere (the reference's reset value)."""
        for t, n in ((C, "C"), (u, "u"), (v, "v"), (p, "p"), (rho, "rho")):
            self._field(t, n)
        if T is not None:
            self._field(T, "T")
        out = (ct.c_double * 8)()
        self._stream()
        L.check(self._lib.covof_history_d(self._h, _ptr(C), _ptr(u), _ptr(v), _ptr(p), _ptr(rho), _ptr(T), out),
                "covof_history_d")
        return list(out)

    def cell_output(self, u, v):
        """Divergence and cell-centred velocity of vtk_write_step, (ny, nx)."""
        self._field(u, "u")
        self._field(v, "v")
        div, uc, vc = (self.new_work() for _ in range(3))
        self._stream()
        L.check(self._lib.covof_cell_output_d(self._h, _ptr(u), _ptr(v), _ptr(div), _ptr(uc), _ptr(vc)),
                "covof_cell_output_d")
        return div, uc, vc
