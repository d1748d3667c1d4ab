!> The reference's only extension point routed to the GPU: a procedure with
!> the interface of custom_precond_proc (src/module_linear_solvers.f90:22-34),
!> for custom_precond_hook (:36), which apply_preconditioner calls when the
!> solve's preconditioner is PRECOND_CUSTOM (:306-312).
!>
!> This file is compiled with the reference's modules (it uses their types for
!> the hook's dummy arguments) and linked with libcovof_hip and covof_hip.o.
!> Use, in the reference's time loop, around a solve_linear_system
!> (src/module_linear_solvers.f90:1167) of the momentum block with
!> cfg%preconditioner = PRECOND_CUSTOM:
!>
!>   st = covof_momentum_set_fields(ctx, mcfg, u, v, u_old, v_old, u_nm1, v_nm1,
!>                                  rho_u, rho_v, mu, mu_node, fx, fy, dt)
!>   call covof_precond_hook_install(ctx, COVOF_PRECOND_SSOR)
!>   call solve_linear_system(ns, mesh, cfg, fields_old, dt, b, x, iters)
!>   if (covof_precond_status() /= COVOF_OK) ...   ! a call fell back to dst = src
!>
!> The fields are those the block's assemble_preconditioner_blocks reads
!> (fields_old, halo layout); the GPU then applies JACOBI, SSOR or
!> BLOCK_TRIANGULAR to src on the device, the same arithmetic as the
!> reference's own apply_preconditioner branches, so with SSOR the
!> PRECOND_CUSTOM solve takes the PRECOND_SSOR solve's iterates.
module covof_precond_hook
  use, intrinsic :: iso_c_binding, only: c_ptr, c_int
  use module_precision, only: dp
  use module_mesh, only: mesh_2d_t
  use module_config, only: config_t
  use module_navier_stokes, only: navier_stokes_t
  use module_linear_solvers, only: custom_precond_hook
  use covof_hip, only: covof_precond_select, covof_precond_apply
  implicit none
  private
  public :: covof_gpu_precond, covof_precond_hook_install, covof_precond_hook_remove

contains

  !> custom_precond_proc: dst = M^-1 src on the GPU (ns, mesh and cfg are the
  !> reference's arguments; the operator comes from the fields bound to the
  !> selected context).
  subroutine covof_gpu_precond(ns, mesh, cfg, src, dst)
    type(navier_stokes_t), intent(in) :: ns
    type(mesh_2d_t), intent(in) :: mesh
    type(config_t), intent(in) :: cfg
    real(dp), intent(in) :: src(:)
    real(dp), intent(out) :: dst(:)
    call covof_precond_apply(src, dst)
  end subroutine covof_gpu_precond

  !> Points custom_precond_hook at covof_gpu_precond with context ctx and
  !> GPU preconditioner precond.
  subroutine covof_precond_hook_install(ctx, precond)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: precond
    call covof_precond_select(ctx, precond)
    custom_precond_hook => covof_gpu_precond
  end subroutine covof_precond_hook_install

  !> Restores the reference's default (no hook: dst = src).
  subroutine covof_precond_hook_remove()
    custom_precond_hook => null()
  end subroutine covof_precond_hook_remove
end module covof_precond_hook
