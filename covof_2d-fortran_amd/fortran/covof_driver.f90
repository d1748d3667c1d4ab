!> Minimal Fortran driver of libcovof_hip through the covof_hip ISO_C_BINDING
!> module, with the call shapes of the reference's time loop:
!>   vof  : advect_vof(fields, mesh, cfg, dt) with WEYMOUTH_YUE + ELVIRA
!>          (src/module_interface_vof.f90:792) on a single vortex, nsteps steps;
!>   sp   : the Poisson block and the velocity correction of
!>          sim_solve_standard_projection (src/module_simulation.f90:2477-2654)
!>          with the reference's CG + SSOR, tol 1e-7, 1000 iterations, on a
!>          walled box holding a heavy disc (rho 1000 in 1);
!>   csf  : the tail of advect_vof (src/module_interface_vof.f90:810-824) on a
!>          walled box holding a disc: smooth_svof_helmholtz (filter width
!>          1.5), vof_smooth_diffusion (NSTEPS steps) and
!>          compute_curvature_and_csf with the normals from C_tilde, sigma 0.07;
!>   pred : the momentum block of sim_solve_standard_projection
!>          (src/module_simulation.f90:2454-2469) on a lid-driven box holding a
!>          heavy viscous disc (rho 1000 / 100, mu 10 / 1), BDF2 with an n-1
!>          state, TVD superbee convection, gravity, BiCGSTAB(2) + Jacobi to 1e-10;
!>   mono : the monolithic branch of sim_run (src/module_simulation.f90:550-583)
!>          on the same box: u, v, p coupled, BiCGSTAB(2) + block-triangular
!>          Schur-SSOR to 1e-10;
!>   hook : the body of the PRECOND_CUSTOM hook (covof_precond_hook.f90) on
!>          the pred box's fields, bound from host arrays: dst = M^-1 src with
!>          the GPU's SSOR, src(k) = sin(0.37 k) + 0.1.
!>
!> usage: covof_driver vof|sp|csf|pred|mono|hook N NSTEPS OUTFILE
!>
!> OUTFILE (unformatted stream, native endianness):
!>   int32 mode (1 vof, 2 sp), nx, ny, nsteps, counter (vof: final
!>   vof_step_count; sp: Poisson iterations), int32 status;
!>   float64 dt, relres (sp; 0 for vof);
!>   vof: C0, u, v, C            (each (0:nx+1, 0:ny+1))
!>   sp : u*, v*, rho_u, rho_v, p, u, v (each (0:nx+1, 0:ny+1))
!>   csf: C, C_s, C_tilde, curvature, fx_sigma, fy_sigma
!>   pred: u0, v0, u_nm1, v_nm1, rho_u, rho_v, mu, mu_node, u*, v*
!>   mono: u0, v0, u_nm1, v_nm1, rho_u, rho_v, mu, mu_node, u, v, p
!>   hook: u0, v0, u_nm1, v_nm1, rho_u, rho_v, mu, mu_node, then src and dst
!>         (counter entries each; counter = the momentum system's size)
program covof_driver
  use iso_c_binding
  use covof_hip
  implicit none
  integer, parameter :: dp = c_double
  character(len=16) :: mode, arg
  character(len=512) :: outfile
  integer :: n, nsteps, i, j, st, unit, counter
  integer(c_int) :: bc(4), step_count, iters
  type(c_ptr) :: ctx
  real(dp) :: dx, dy, dt, relres, umax
  real(dp), allocatable :: C(:, :), C0(:, :), u(:, :), v(:, :), psi(:, :)
  real(dp), allocatable :: rho(:, :), rho_u(:, :), rho_v(:, :), p(:, :), us(:, :), vs(:, :)
  real(dp), allocatable :: Cs(:, :), Ct(:, :), curv(:, :), fx(:, :), fy(:, :)
  real(dp), allocatable :: mu(:, :), mun(:, :), u0(:, :), v0(:, :), um(:, :), vm(:, :)
  real(dp), allocatable :: src(:), dst(:)
  integer(c_int) :: nm, unx, vny
  type(covof_momentum_cfg) :: mcfg
  real(dp), parameter :: pi = 3.14159265358979323846_dp

  if (command_argument_count() < 4) then
     print '(a)', "usage: covof_driver vof|sp|csf|pred|mono|hook N NSTEPS OUTFILE"
     stop 2
  end if
  call get_command_argument(1, mode)
  call get_command_argument(2, arg); read (arg, *) n
  call get_command_argument(3, arg); read (arg, *) nsteps
  call get_command_argument(4, outfile)
  dx = 1.0_dp / n
  dy = 1.0_dp / n
  allocate(C(0:n + 1, 0:n + 1), C0(0:n + 1, 0:n + 1), u(0:n + 1, 0:n + 1), v(0:n + 1, 0:n + 1))
  allocate(psi(0:n + 2, 0:n + 2))
  relres = 0.0_dp
  counter = 0

  select case (trim(mode))
  case ("vof")
     bc = COVOF_BC_PERIODIC
     st = covof_ctx_create(int(n, c_int), int(n, c_int), 1.0_dp, 1.0_dp, bc, ctx)
     call check(st, "covof_ctx_create")
     call disc(C, n, 0.5_dp, 0.75_dp, 0.15_dp)
     call wrap(C, n)
     ! nodal stream function of the single vortex, psi = sin^2(pi x) sin^2(pi y) / pi
     do j = 0, n + 2
        do i = 0, n + 2
           psi(i, j) = sin(pi * (i - 1) * dx)**2 * sin(pi * (j - 1) * dy)**2 / pi
        end do
     end do
     do j = 0, n + 1
        do i = 0, n + 1
           u(i, j) = (psi(i, j + 1) - psi(i, j)) / dy
           v(i, j) = -(psi(i + 1, j) - psi(i, j)) / dx
        end do
     end do
     umax = max(maxval(abs(u)), maxval(abs(v)))
     dt = 0.25_dp * dx / umax
     C0 = C
     step_count = 0
     st = covof_advect_vof(ctx, C, u, v, dt, "WEYMOUTH_YUE", "ELVIRA", step_count, int(nsteps, c_int))
     call check(st, "covof_advect_vof")
     counter = step_count
     open (newunit=unit, file=trim(outfile), access="stream", form="unformatted", status="replace")
     write (unit) 1_c_int, int(n, c_int), int(n, c_int), int(nsteps, c_int), int(counter, c_int), int(st, c_int)
     write (unit) dt, relres
     write (unit) C0, u, v, C
     close (unit)
     print '(a,i0,a,es23.16)', "vof steps ", counter, " volume ", sum(C(1:n, 1:n)) * dx * dy
  case ("sp")
     bc = COVOF_BC_FREESLIP
     st = covof_ctx_create(int(n, c_int), int(n, c_int), 1.0_dp, 1.0_dp, bc, ctx)
     call check(st, "covof_ctx_create")
     allocate(rho(0:n + 1, 0:n + 1), rho_u(0:n + 1, 0:n + 1), rho_v(0:n + 1, 0:n + 1), p(0:n + 1, 0:n + 1))
     allocate(us(0:n + 1, 0:n + 1), vs(0:n + 1, 0:n + 1))
     call disc(C, n, 0.45_dp, 0.55_dp, 0.2_dp)
     ! zero-gradient halo, rho = C rho1 + (1 - C) rho2, arithmetic face means
     C(0, :) = C(1, :); C(n + 1, :) = C(n, :); C(:, 0) = C(:, 1); C(:, n + 1) = C(:, n)
     rho = C * 1000.0_dp + (1.0_dp - C) * 1.0_dp
     rho_u = 1.0_dp
     rho_v = 1.0_dp
     do j = 0, n + 1
        do i = 1, n + 1
           rho_u(i, j) = 0.5_dp * (rho(i - 1, j) + rho(i, j))
        end do
     end do
     do j = 1, n + 1
        do i = 0, n + 1
           rho_v(i, j) = 0.5_dp * (rho(i, j - 1) + rho(i, j))
        end do
     end do
     ! a predicted velocity with divergence: smooth modes, zero on the walls
     do j = 0, n + 1
        do i = 0, n + 1
           us(i, j) = sin(3.0_dp * pi * (i - 1) * dx) * cos(2.0_dp * pi * (j - 0.5_dp) * dy) &
                + 0.3_dp * sin(pi * (i - 1) * dx) * sin(5.0_dp * pi * (j - 0.5_dp) * dy)
           vs(i, j) = 0.5_dp * cos(pi * (i - 0.5_dp) * dx) * sin(4.0_dp * pi * (j - 1) * dy)
        end do
     end do
     us(1, :) = 0.0_dp; us(n + 1, :) = 0.0_dp; vs(:, 1) = 0.0_dp; vs(:, n + 1) = 0.0_dp
     u = us
     v = vs
     p = 0.0_dp
     dt = 1.0e-3_dp
     st = covof_sp_poisson(ctx, u, v, rho_u, rho_v, p, dt, COVOF_LINEAR_CG, COVOF_PRECOND_SSOR, 1.0e-7_dp, &
          1000_c_int, COVOF_PRECISION_FP64, iters, relres)
     if (st /= COVOF_OK .and. st /= COVOF_NOT_CONVERGED) call check(st, "covof_sp_poisson")
     counter = iters
     st = covof_sp_correct(ctx, p, rho_u, rho_v, dt, u, v)
     call check(st, "covof_sp_correct")
     open (newunit=unit, file=trim(outfile), access="stream", form="unformatted", status="replace")
     write (unit) 2_c_int, int(n, c_int), int(n, c_int), int(nsteps, c_int), int(counter, c_int), int(st, c_int)
     write (unit) dt, relres
     write (unit) us, vs, rho_u, rho_v, p, u, v
     close (unit)
     print '(a,i0,a,es10.3)', "sp poisson iterations ", counter, " relres ", relres
  case ("csf")
     bc = COVOF_BC_FREESLIP
     st = covof_ctx_create(int(n, c_int), int(n, c_int), 1.0_dp, 1.0_dp, bc, ctx)
     call check(st, "covof_ctx_create")
     allocate(Cs(0:n + 1, 0:n + 1), Ct(0:n + 1, 0:n + 1), curv(0:n + 1, 0:n + 1))
     allocate(fx(0:n + 1, 0:n + 1), fy(0:n + 1, 0:n + 1))
     call disc(C, n, 0.45_dp, 0.55_dp, 0.27_dp)
     C(0, :) = C(1, :); C(n + 1, :) = C(n, :); C(:, 0) = C(:, 1); C(:, n + 1) = C(:, n)
     fx = 0.0_dp
     fy = 0.0_dp
     st = covof_svof_smooth(ctx, C, 1.5_dp, Cs)
     call check(st, "covof_svof_smooth")
     st = covof_smooth_diffusion(ctx, C, int(nsteps, c_int), Ct)
     call check(st, "covof_smooth_diffusion")
     st = covof_curvature_csf(ctx, Ct, C, 1_c_int, 0.07_dp, curv, fx, fy)
     call check(st, "covof_curvature_csf")
     dt = 0.0_dp
     open (newunit=unit, file=trim(outfile), access="stream", form="unformatted", status="replace")
     write (unit) 3_c_int, int(n, c_int), int(n, c_int), int(nsteps, c_int), int(counter, c_int), int(st, c_int)
     write (unit) dt, relres
     write (unit) C, Cs, Ct, curv, fx, fy
     close (unit)
     print '(a,es23.16)', "csf max |curvature| ", maxval(abs(curv))
  case ("pred", "mono", "hook")
     bc = COVOF_BC_DIRICHLET
     st = covof_ctx_create(int(n, c_int), int(n, c_int), 1.0_dp, 1.0_dp, bc, ctx)
     call check(st, "covof_ctx_create")
     allocate(rho(0:n + 1, 0:n + 1), rho_u(0:n + 1, 0:n + 1), rho_v(0:n + 1, 0:n + 1))
     allocate(mu(0:n + 1, 0:n + 1), mun(0:n + 1, 0:n + 1), u0(0:n + 1, 0:n + 1), v0(0:n + 1, 0:n + 1))
     allocate(um(0:n + 1, 0:n + 1), vm(0:n + 1, 0:n + 1), fx(0:n + 1, 0:n + 1), fy(0:n + 1, 0:n + 1))
     call disc(C, n, 0.5_dp, 0.4_dp, 0.2_dp)
     C(0, :) = C(1, :); C(n + 1, :) = C(n, :); C(:, 0) = C(:, 1); C(:, n + 1) = C(:, n)
     rho = C * 1000.0_dp + (1.0_dp - C) * 100.0_dp
     mu = C * 10.0_dp + (1.0_dp - C) * 1.0_dp
     rho_u = rho
     rho_v = rho
     mun = mu
     do j = 0, n + 1
        do i = 1, n + 1
           rho_u(i, j) = 0.5_dp * (rho(i - 1, j) + rho(i, j))
        end do
     end do
     do j = 1, n + 1
        do i = 0, n + 1
           rho_v(i, j) = 0.5_dp * (rho(i, j - 1) + rho(i, j))
        end do
     end do
     do j = 1, n + 1
        do i = 1, n + 1
           mun(i, j) = 0.25_dp * (mu(i - 1, j - 1) + mu(i, j - 1) + mu(i - 1, j) + mu(i, j))
        end do
     end do
     do j = 0, n + 1
        do i = 0, n + 1
           u0(i, j) = 0.2_dp * sin(2.0_dp * pi * (i - 1) * dx) * cos(pi * (j - 0.5_dp) * dy)
           v0(i, j) = -0.1_dp * cos(pi * (i - 0.5_dp) * dx) * sin(2.0_dp * pi * (j - 1) * dy)
        end do
     end do
     um = 0.9_dp * u0
     vm = 0.9_dp * v0
     u = u0
     v = v0
     fx = 0.0_dp
     fy = 0.0_dp
     dt = 1.0e-3_dp
     mcfg%time_scheme = COVOF_TIME_SCHEME_GEAR
     mcfg%has_nm1 = 1
     mcfg%convective_scheme = COVOF_CONV_TVD_SUPERBEE
     mcfg%gy = -0.98_dp
     mcfg%bc_u(4) = 1.0_dp
     if (trim(mode) == "hook") then
        st = covof_momentum_set_fields(ctx, mcfg, u0, v0, u0, v0, um, vm, rho_u, rho_v, mu, mun, fx, fy, dt)
        call check(st, "covof_momentum_set_fields")
        st = covof_momentum_size(ctx, COVOF_SOLVER_STANDARD_PROJECTION, nm, unx, vny)
        call check(st, "covof_momentum_size")
        allocate(src(nm), dst(nm))
        do i = 1, nm
           src(i) = sin(0.37_dp * i) + 0.1_dp
        end do
        call covof_precond_select(ctx, COVOF_PRECOND_SSOR)
        call covof_precond_apply(src, dst)
        st = covof_precond_status()
        call check(st, "covof_precond_apply")
        open (newunit=unit, file=trim(outfile), access="stream", form="unformatted", status="replace")
        write (unit) 6_c_int, int(n, c_int), int(n, c_int), int(nsteps, c_int), nm, int(st, c_int)
        write (unit) dt, 0.0_dp
        write (unit) u0, v0, um, vm, rho_u, rho_v, mu, mun, src, dst
        close (unit)
        print '(a,i0,a,es23.16)', "hook: ", nm, " unknowns, max |dst| ", maxval(abs(dst))
        st = covof_ctx_destroy(ctx)
        stop
     else if (trim(mode) == "pred") then
        st = covof_sp_predictor(ctx, mcfg, u, v, u0, v0, um, vm, rho_u, rho_v, mu, mun, fx, fy, dt, &
             COVOF_LINEAR_BICGSTAB2, COVOF_PRECOND_JACOBI, 1.0e-10_dp, 300_c_int, iters, relres)
        call check(st, "covof_sp_predictor")
     else
        allocate(p(0:n + 1, 0:n + 1))
        p = 0.0_dp
        mcfg%solver_type = COVOF_SOLVER_MONOLITHIC
        st = covof_monolithic_step(ctx, mcfg, u, v, p, u0, v0, um, vm, rho_u, rho_v, mu, mun, fx, fy, dt, &
             COVOF_LINEAR_BICGSTAB2, COVOF_PRECOND_BLOCK_TRIANGULAR, 1.0e-10_dp, 600_c_int, iters, relres)
        call check(st, "covof_monolithic_step")
     end if
     counter = iters
     open (newunit=unit, file=trim(outfile), access="stream", form="unformatted", status="replace")
     write (unit) merge(4_c_int, 5_c_int, trim(mode) == "pred"), int(n, c_int), int(n, c_int), &
          int(nsteps, c_int), int(counter, c_int), int(st, c_int)
     write (unit) dt, relres
     write (unit) u0, v0, um, vm, rho_u, rho_v, mu, mun, u, v
     if (trim(mode) == "mono") write (unit) p
     close (unit)
     print '(a,a,a,i0,a,es10.3)', "solve (", trim(mode), ") iterations ", counter, " relres ", relres
  case default
     print '(a)', "usage: covof_driver vof|sp|csf|pred|mono|hook N NSTEPS OUTFILE"
     stop 2
  end select
  st = covof_ctx_destroy(ctx)

contains

  subroutine check(status, where)
    integer, intent(in) :: status
    character(len=*), intent(in) :: where
    if (status /= COVOF_OK) then
       print '(a,a,i0,a,a)', where, ": status ", status, " ", covof_last_error()
       stop 1
    end if
  end subroutine check

  !> Volume fraction of a disc by 16 x 16 sub-cell sampling.
  subroutine disc(F, n, xc, yc, r)
    real(dp), intent(out) :: F(0:, 0:)
    integer, intent(in) :: n
    real(dp), intent(in) :: xc, yc, r
    integer :: a, b, ii, jj, hits
    real(dp) :: x, y, h
    h = 1.0_dp / n
    F = 0.0_dp
    do jj = 1, n
       do ii = 1, n
          hits = 0
          do b = 0, 15
             do a = 0, 15
                x = (ii - 1 + (a + 0.5_dp) / 16.0_dp) * h
                y = (jj - 1 + (b + 0.5_dp) / 16.0_dp) * h
                if ((x - xc)**2 + (y - yc)**2 <= r * r) hits = hits + 1
             end do
          end do
          F(ii, jj) = hits / 256.0_dp
       end do
    end do
  end subroutine disc

  subroutine wrap(F, n)
    real(dp), intent(inout) :: F(0:, 0:)
    integer, intent(in) :: n
    F(0, 1:n) = F(n, 1:n); F(n + 1, 1:n) = F(1, 1:n)
    F(:, 0) = F(:, n); F(:, n + 1) = F(:, 1)
  end subroutine wrap
end program covof_driver
