!> ISO_C_BINDING interface of libcovof_hip (include/covof_hip.h) for the
!> reference Fortran driver.
!>
!> This is the module a COVOF build `use`s to route its per-timestep hot path
!> to the GPU:
!>   advect_vof(fields, mesh, cfg, dt)        src/module_interface_vof.f90:792
!>     -> covof_vof_advect(ctx, fields%C, fields%u, fields%v, dt, ...)
!>   reconstruct_interface(method, mesh, C, segments)
!>                                            src/module_interface_reconstruction.f90:580
!>     -> covof_reconstruct_interface_d (device arrays)
!>   the Poisson block of sim_solve_standard_projection
!>                                            src/module_simulation.f90:2477-2588
!>     -> covof_sp_poisson(ctx, u, v, rho_u, rho_v, p, dt, ...)
!>   the velocity correction                  src/module_simulation.f90:2590-2654
!>     -> covof_sp_correct(ctx, p, rho_u, rho_v, dt, u, v)
!>   the momentum block (velocity predictor) of sim_solve_standard_projection
!>                                            src/module_simulation.f90:2454-2469
!>     -> covof_sp_predictor(ctx, mcfg, u, v, u_old, v_old, u_nm1, v_nm1, rho_u,
!>            rho_v, mu, mu_node, fx_sigma, fy_sigma, dt, linear_solver,
!>            precond_type, tol, max_iter, iters, relres)
!>   the monolithic branch of sim_run          src/module_simulation.f90:550-583
!>     -> covof_monolithic_step(ctx, mcfg, u, v, p, u_old, ..., fy_sigma, dt,
!>            linear_solver, precond_type, tol, max_iter, iters, relres)
!>   the tail of advect_vof                   src/module_interface_vof.f90:810-824
!>     smooth_svof_helmholtz   -> covof_svof_smooth(ctx, C, width, C_s)
!>     vof_smooth_diffusion    -> covof_smooth_diffusion(ctx, C, steps, C_tilde)
!>     compute_curvature_and_csf -> covof_curvature_csf(ctx, G, C, enable, sigma,
!>                                     curvature, fx_sigma, fy_sigma)
!>
!> Arrays are the reference's own: real(dp) F(0:nx+1, 0:ny+1), column-major,
!> i fastest (src/module_fields.f90:129-169), passed by reference.  Selector
!> values are those of src/module_constants.f90.  Every function returns an
!> integer status (COVOF_OK = 0).
module covof_hip
  use iso_c_binding
  implicit none
  private :: c_str, hook_ctx, hook_precond, hook_status

  integer(c_int), parameter :: COVOF_OK = 0, COVOF_NOT_CONVERGED = 1, COVOF_BREAKDOWN = 2, &
       COVOF_NONFINITE = 3, COVOF_BAD_ARGUMENT = 4, COVOF_HIP_ERROR = 5
  integer(c_int), parameter :: COVOF_BC_DIRICHLET = 1, COVOF_BC_NEUMANN = 2, COVOF_BC_PERIODIC = 3, &
       COVOF_BC_FREESLIP = 4
  integer(c_int), parameter :: COVOF_LINEAR_BICGSTAB2 = 1, COVOF_LINEAR_CG = 3
  integer(c_int), parameter :: COVOF_PRECOND_NONE = 0, COVOF_PRECOND_JACOBI = 1, COVOF_PRECOND_SSOR = 2, &
       COVOF_PRECOND_BLOCK_TRIANGULAR = 4
  integer(c_int), parameter :: COVOF_SOLVER_MONOLITHIC = 2, COVOF_SOLVER_STANDARD_PROJECTION = 3
  !> parallel preconditioners of the GPU library (no reference counterpart)
  integer(c_int), parameter :: COVOF_PRECOND_SGS_RB = 16, COVOF_PRECOND_MG = 17
  integer(c_int), parameter :: COVOF_PRECISION_FP64 = 0, COVOF_PRECISION_MIXED = 1
  !> summation order of the FP64 inner products (covof_ctx_set_reduction_order)
  integer(c_int), parameter :: COVOF_REDUCTION_TREE = 0, COVOF_REDUCTION_REFERENCE = 1
  integer(c_int), parameter :: COVOF_TIME_SCHEME_EULER = 1, COVOF_TIME_SCHEME_GEAR = 2
  integer(c_int), parameter :: COVOF_CONV_UPWIND = 1, COVOF_CONV_CENTERED = 2, COVOF_CONV_TVD_SUPERBEE = 3
  integer(c_int), parameter :: COVOF_CONV_LIN_PICARD = 1, COVOF_CONV_LIN_ADAMS_BASHFORTH = 2, &
       COVOF_CONV_LIN_NEWTON = 3

  !> covof_momentum_cfg: the config_t fields the momentum block reads
  !> (time_scheme, fields%has_nm1, enable_convection, convective_scheme,
  !> convective_linearization, gx, gy, bc_<side>%u / %v of left, right, bottom, top,
  !> solver_type: 3 STANDARD_PROJECTION, 2 MONOLITHIC)
  type, bind(C) :: covof_momentum_cfg
     integer(c_int) :: time_scheme = 1, has_nm1 = 0, enable_convection = 1, convective_scheme = 1, &
          convective_linearization = 1
     real(c_double) :: gx = 0.0_c_double, gy = 0.0_c_double
     real(c_double) :: bc_u(4) = 0.0_c_double, bc_v(4) = 0.0_c_double
     integer(c_int) :: solver_type = 3
  end type covof_momentum_cfg

  interface
     integer(c_int) function covof_ctx_create(nx, ny, lx, ly, bc, ctx) bind(C, name="covof_ctx_create")
       import :: c_int, c_double, c_ptr
       integer(c_int), value :: nx, ny
       real(c_double), value :: lx, ly
       integer(c_int), intent(in) :: bc(4)
       type(c_ptr), intent(out) :: ctx
     end function covof_ctx_create

     integer(c_int) function covof_ctx_destroy(ctx) bind(C, name="covof_ctx_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function covof_ctx_destroy

     integer(c_int) function covof_ctx_synchronize(ctx) bind(C, name="covof_ctx_synchronize")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function covof_ctx_synchronize

     integer(c_int) function covof_ctx_set_reduction_order(ctx, order) bind(C, name="covof_ctx_set_reduction_order")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: order
     end function covof_ctx_set_reduction_order

     type(c_ptr) function covof_last_error_c() bind(C, name="covof_last_error")
       import :: c_ptr
     end function covof_last_error_c

     integer(c_int) function covof_vof_advect_c(ctx, C, u, v, dt, adv, rec, step_count, nsteps, two_phase, &
          rho1, rho2, mu1, mu2, rho, rho_u, rho_v, mu, mu_node) bind(C, name="covof_vof_advect")
       import :: c_int, c_double, c_ptr, c_char
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: C(*)
       real(c_double), intent(in) :: u(*), v(*)
       real(c_double), value :: dt
       character(kind=c_char), intent(in) :: adv(*), rec(*)
       integer(c_int), intent(inout) :: step_count
       integer(c_int), value :: nsteps, two_phase
       real(c_double), value :: rho1, rho2, mu1, mu2
       type(c_ptr), value :: rho, rho_u, rho_v, mu, mu_node
     end function covof_vof_advect_c

     integer(c_int) function covof_sp_poisson(ctx, u, v, rho_u, rho_v, p, dt, lin_solver, precond, tol, maxit, &
          precision, iters, relres) bind(C, name="covof_sp_poisson")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: u(*), v(*), rho_u(*), rho_v(*)
       real(c_double), intent(inout) :: p(*)
       real(c_double), value :: dt, tol
       integer(c_int), value :: lin_solver, precond, maxit, precision
       integer(c_int), intent(out) :: iters
       real(c_double), intent(out) :: relres
     end function covof_sp_poisson

     integer(c_int) function covof_sp_correct(ctx, p, rho_u, rho_v, dt, u, v) bind(C, name="covof_sp_correct")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: p(*), rho_u(*), rho_v(*)
       real(c_double), value :: dt
       real(c_double), intent(inout) :: u(*), v(*)
     end function covof_sp_correct

     integer(c_int) function covof_sp_predictor(ctx, mcfg, u, v, u_old, v_old, u_nm1, v_nm1, rho_u, rho_v, &
          mu, mu_node, fx_sigma, fy_sigma, dt, lin_solver, precond, tol, maxit, iters, relres) &
          bind(C, name="covof_sp_predictor")
       import :: c_int, c_double, c_ptr, covof_momentum_cfg
       type(c_ptr), value :: ctx
       type(covof_momentum_cfg), intent(in) :: mcfg
       real(c_double), intent(inout) :: u(*), v(*)
       real(c_double), intent(in) :: u_old(*), v_old(*), u_nm1(*), v_nm1(*), rho_u(*), rho_v(*), mu(*), &
            mu_node(*), fx_sigma(*), fy_sigma(*)
       real(c_double), value :: dt, tol
       integer(c_int), value :: lin_solver, precond, maxit
       integer(c_int), intent(out) :: iters
       real(c_double), intent(out) :: relres
     end function covof_sp_predictor

     integer(c_int) function covof_momentum_set_fields(ctx, mcfg, u, v, u_old, v_old, u_nm1, v_nm1, rho_u, &
          rho_v, mu, mu_node, fx_sigma, fy_sigma, dt) bind(C, name="covof_momentum_set_fields")
       import :: c_int, c_double, c_ptr, covof_momentum_cfg
       type(c_ptr), value :: ctx
       type(covof_momentum_cfg), intent(in) :: mcfg
       real(c_double), intent(in) :: u(*), v(*), u_old(*), v_old(*), u_nm1(*), v_nm1(*), rho_u(*), rho_v(*), &
            mu(*), mu_node(*), fx_sigma(*), fy_sigma(*)
       real(c_double), value :: dt
     end function covof_momentum_set_fields

     integer(c_int) function covof_momentum_size(ctx, solver_type, n, u_nx, v_ny) bind(C, name="covof_momentum_size")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: solver_type
       integer(c_int), intent(out) :: n, u_nx, v_ny
     end function covof_momentum_size

     integer(c_int) function covof_momentum_precond(ctx, precond, n, r, z) bind(C, name="covof_momentum_precond")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: precond, n
       real(c_double), intent(in) :: r(*)
       real(c_double), intent(out) :: z(*)
     end function covof_momentum_precond

     integer(c_int) function covof_monolithic_step(ctx, mcfg, u, v, p, u_old, v_old, u_nm1, v_nm1, rho_u, &
          rho_v, mu, mu_node, fx_sigma, fy_sigma, dt, lin_solver, precond, tol, maxit, iters, relres) &
          bind(C, name="covof_monolithic_step")
       import :: c_int, c_double, c_ptr, covof_momentum_cfg
       type(c_ptr), value :: ctx
       type(covof_momentum_cfg), intent(in) :: mcfg
       real(c_double), intent(inout) :: u(*), v(*), p(*)
       real(c_double), intent(in) :: u_old(*), v_old(*), u_nm1(*), v_nm1(*), rho_u(*), rho_v(*), mu(*), &
            mu_node(*), fx_sigma(*), fy_sigma(*)
       real(c_double), value :: dt, tol
       integer(c_int), value :: lin_solver, precond, maxit
       integer(c_int), intent(out) :: iters
       real(c_double), intent(out) :: relres
     end function covof_monolithic_step

     integer(c_int) function covof_curvature_csf(ctx, G, C, enable, sigma, curvature, fx_sigma, fy_sigma) &
          bind(C, name="covof_curvature_csf")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: G(*), C(*)
       integer(c_int), value :: enable
       real(c_double), value :: sigma
       real(c_double), intent(inout) :: curvature(*), fx_sigma(*), fy_sigma(*)
     end function covof_curvature_csf

     integer(c_int) function covof_smooth_diffusion(ctx, C, steps, C_tilde) bind(C, name="covof_smooth_diffusion")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: C(*)
       integer(c_int), value :: steps
       real(c_double), intent(inout) :: C_tilde(*)
     end function covof_smooth_diffusion

     integer(c_int) function covof_svof_smooth(ctx, C, filter_width, C_s) bind(C, name="covof_svof_smooth")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: C(*)
       real(c_double), value :: filter_width
       real(c_double), intent(inout) :: C_s(*)
     end function covof_svof_smooth

     ! device-pointer entry points (arrays already resident in HBM)
     integer(c_int) function covof_vof_advect_d(ctx, C, u, v, dt, adv, rec, step_count, nsteps, two_phase, &
          rho1, rho2, mu1, mu2, rho, rho_u, rho_v, mu, mu_node) bind(C, name="covof_vof_advect_d")
       import :: c_int, c_double, c_ptr, c_char
       type(c_ptr), value :: ctx, C, u, v
       real(c_double), value :: dt
       character(kind=c_char), intent(in) :: adv(*), rec(*)
       integer(c_int), intent(inout) :: step_count
       integer(c_int), value :: nsteps, two_phase
       real(c_double), value :: rho1, rho2, mu1, mu2
       type(c_ptr), value :: rho, rho_u, rho_v, mu, mu_node
     end function covof_vof_advect_d

     integer(c_int) function covof_reconstruct_interface_d(ctx, method, C, seg_nx, seg_ny, seg_alpha) &
          bind(C, name="covof_reconstruct_interface_d")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: method(*)
       type(c_ptr), value :: C, seg_nx, seg_ny, seg_alpha
     end function covof_reconstruct_interface_d

     integer(c_int) function covof_curvature_csf_d(ctx, G, C, enable, sigma, curvature, fx_sigma, fy_sigma) &
          bind(C, name="covof_curvature_csf_d")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx, G, C
       integer(c_int), value :: enable
       real(c_double), value :: sigma
       type(c_ptr), value :: curvature, fx_sigma, fy_sigma
     end function covof_curvature_csf_d

     integer(c_int) function covof_smooth_diffusion_d(ctx, C, steps, C_tilde) bind(C, name="covof_smooth_diffusion_d")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, C
       integer(c_int), value :: steps
       type(c_ptr), value :: C_tilde
     end function covof_smooth_diffusion_d

     integer(c_int) function covof_svof_smooth_d(ctx, C, filter_width, C_s) bind(C, name="covof_svof_smooth_d")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx, C
       real(c_double), value :: filter_width
       type(c_ptr), value :: C_s
     end function covof_svof_smooth_d

     integer(c_int) function covof_sp_poisson_d(ctx, u, v, rho_u, rho_v, p, dt, lin_solver, precond, tol, maxit, &
          precision, iters, relres) bind(C, name="covof_sp_poisson_d")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx, u, v, rho_u, rho_v, p
       real(c_double), value :: dt, tol
       integer(c_int), value :: lin_solver, precond, maxit, precision
       integer(c_int), intent(out) :: iters
       real(c_double), intent(out) :: relres
     end function covof_sp_poisson_d
  end interface

  ! The GPU preconditioner behind the PRECOND_CUSTOM hook (covof_precond_hook.f90):
  ! the context holding the bound fields, the preconditioner, and the status
  ! of the last application that failed (0: none).
  type(c_ptr), save :: hook_ctx = c_null_ptr
  integer(c_int), save :: hook_precond = COVOF_PRECOND_SSOR
  integer(c_int), save :: hook_status = COVOF_OK

contains

  !> Selects the context and the preconditioner (NONE, JACOBI, SSOR or
  !> BLOCK_TRIANGULAR of apply_preconditioner, src/module_linear_solvers.f90:
  !> 48-316) that covof_precond_apply uses; clears the failure status.
  subroutine covof_precond_select(ctx, precond)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: precond
    hook_ctx = ctx
    hook_precond = precond
    hook_status = COVOF_OK
  end subroutine covof_precond_select

  !> dst = M^-1 src on the GPU with the selected context's bound fields (the
  !> body of the custom_precond_proc hook).  If the call fails, dst = src,
  !> what apply_preconditioner does for PRECOND_CUSTOM with no hook
  !> associated (src/module_linear_solvers.f90:310-311), and the status is
  !> kept for covof_precond_status.
  subroutine covof_precond_apply(src, dst)
    real(c_double), intent(in) :: src(:)
    real(c_double), intent(out) :: dst(:)
    integer(c_int) :: st
    st = covof_momentum_precond(hook_ctx, hook_precond, int(size(src), c_int), src, dst)
    if (st /= COVOF_OK) then
       hook_status = st
       dst = src
    end if
  end subroutine covof_precond_apply

  !> Status of the last failed covof_precond_apply since covof_precond_select.
  integer(c_int) function covof_precond_status()
    covof_precond_status = hook_status
  end function covof_precond_status

  !> trim(s) // NUL for the C string arguments.
  function c_str(s) result(cs)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: cs
    cs = trim(s) // c_null_char
  end function c_str

  !> advect_vof(fields, mesh, cfg, dt) on host arrays C, u, v (0:nx+1, 0:ny+1):
  !> nsteps steps of cfg%advection_method with cfg%reconstruction_method;
  !> step_count is the reference's vof_step_count.  Without the optional
  !> property arrays, update_physical_properties is not run.
  integer(c_int) function covof_advect_vof(ctx, C, u, v, dt, advection_method, reconstruction_method, &
       step_count, nsteps, rho1, rho2, mu1, mu2, rho, rho_u, rho_v, mu, mu_node) result(st)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(inout) :: C(:, :)
    real(c_double), intent(in) :: u(:, :), v(:, :)
    real(c_double), intent(in) :: dt
    character(len=*), intent(in) :: advection_method, reconstruction_method
    integer(c_int), intent(inout) :: step_count
    integer(c_int), intent(in) :: nsteps
    real(c_double), intent(in), optional :: rho1, rho2, mu1, mu2
    real(c_double), intent(inout), target, optional :: rho(:, :), rho_u(:, :), rho_v(:, :), mu(:, :), mu_node(:, :)
    type(c_ptr) :: prho, prho_u, prho_v, pmu, pmu_node
    integer(c_int) :: two_phase
    real(c_double) :: r1, r2, m1, m2
    two_phase = 0
    prho = c_null_ptr; prho_u = c_null_ptr; prho_v = c_null_ptr; pmu = c_null_ptr; pmu_node = c_null_ptr
    r1 = 1.0_c_double; r2 = 1.0_c_double; m1 = 1.0_c_double; m2 = 1.0_c_double
    if (present(rho) .and. present(rho_u) .and. present(rho_v) .and. present(mu) .and. present(mu_node) .and. &
         present(rho1) .and. present(rho2) .and. present(mu1) .and. present(mu2)) then
       two_phase = 1
       r1 = rho1; r2 = rho2; m1 = mu1; m2 = mu2
       prho = c_loc(rho); prho_u = c_loc(rho_u); prho_v = c_loc(rho_v); pmu = c_loc(mu); pmu_node = c_loc(mu_node)
    end if
    st = covof_vof_advect_c(ctx, C, u, v, dt, c_str(advection_method), c_str(reconstruction_method), &
         step_count, nsteps, two_phase, r1, r2, m1, m2, prho, prho_u, prho_v, pmu, pmu_node)
  end function covof_advect_vof

  !> Message of the last failed call.
  function covof_last_error() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    p = covof_last_error_c()
    if (.not. c_associated(p)) then
       msg = ""
       return
    end if
    call c_f_pointer(p, chars, [4096])
    n = 0
    do while (n < 4096)
       if (chars(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(chars(1:n), msg)
  end function covof_last_error
end module covof_hip
