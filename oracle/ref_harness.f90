! Fixture harness over the compiled reference Fortran (test infrastructure only).
!
! This program is original code of this repository.  It is compiled together
! with the reference's own module sources (read in place under the reference
! tree by oracle/Makefile, output only into oracle/_ref/) and calls the
! reference's PUBLIC procedures on inputs read from a binary file, writing the
! reference's outputs bit-for-bit (unformatted stream, little-endian FP64):
!
!   geom    cut_box_area / compute_alpha
!           (src/module_interface_reconstruction.f90:33-147)
!   recon   reconstruct_interface, PLIC or ELVIRA, incl. ghost-normal fix-up
!           (src/module_interface_reconstruction.f90:580-683)
!   advect  advect_vof with WEYMOUTH_YUE, n steps, state after every step
!           (src/module_interface_vof.f90:792-825, :1162-1377), followed by
!           update_physical_properties (:479-616)
!   sp      sim%solve_standard_projection on a small SP case: captures u* by
!           running the same predictor calls first, then the full SP call
!           (src/module_simulation.f90:2424-2666)
!   surface advect_vof with WEYMOUTH_YUE and the optional steps that follow the
!           transport in advect_vof: SVOF smoothing, diffusion smoothing,
!           curvature and CSF forces (src/module_interface_vof.f90:387-476,
!           :627-789, :792-825, :1416-1461); C, C_s, C_tilde, curvature,
!           fx_sigma, fy_sigma after every step
!   momentum the momentum block of the standard projection
!           (src/module_simulation.f90:2454-2469), or the monolithic system
!           (solver type 2): diag_u/diag_v(/diag_p) of
!           assemble_preconditioner_blocks, assemble_rhs, two products of
!           apply_system_matrix, one apply_preconditioner, then pack /
!           solve_linear_system / unpack (and, standard projection, the
!           trailing sync + apply_boundary_conditions); u, v, p and iters
!           (src/module_navier_stokes.f90:231-1101, :1268-1595;
!           src/module_linear_solvers.f90:361-579, :735-838, :1167-1208)
!   output  vtk_writer_t on given fields: init, then write_step and
!           write_diagnostics per snapshot, finalize (src/module_vtk_writer.f90:
!           46-545); the files go to <output.bin>_dir/
!   sim     the reference's own program on a JSON case: sim%init(comm, file)
!           then sim%run() (src/module_simulation.f90:96-351, :416-755); the
!           JSON path is the input file, the outputs go where it says
!   spstep  the body of sim_run's time loop for a two-phase standard-projection
!           run (src/module_simulation.f90:495-559): n-1 / old swaps,
!           advect_vof (WY + properties + CSF), apply_velocity_bc,
!           apply_vof_bc, solve_standard_projection; C, u, v, p and the
!           predictor + Poisson iteration total after every step
!   props   update_physical_properties on given C (and C_s with SVOF
!           smoothing on): C after the physical-corner fix, rho, rho_u, rho_v,
!           mu, mu_node (src/module_interface_vof.f90:479-616)
!   jet     advect_vof with WEYMOUTH_YUE and a JET_INLET top profile, n steps
!           (src/module_interface_vof.f90:792-825; apply_vof_bc,
!           src/module_mpi_comm.f90:636-683): C after every step
!   spobs   one standard-projection step with a solid obstacle (obs_perm 0):
!           the predicted u*, v* (the predictor calls of
!           solve_standard_projection replayed on a copy of the fields), then
!           solve_standard_projection itself; rho_u, rho_v, p, u, v, the
!           iteration total and the predictor's (src/module_simulation.f90:
!           2424-2666, obstacle faces :2491-2528, :2596-2632)
!
! Usage: covof_ref_harness <mode> <input.bin> <output.bin>
program covof_ref_harness
  use module_precision, only: dp
  use module_constants
  use module_config, only: config_t
  use module_mesh, only: mesh_2d_t, init_mesh
  use module_fields, only: flow_fields_t
  use module_interface_reconstruction, only: interface_segment_t, cut_box_area, &
       compute_alpha, reconstruct_interface
  use module_interface_vof, only: advect_vof, update_physical_properties
  use module_navier_stokes, only: navier_stokes_t
  use module_linear_solvers, only: solve_linear_system, apply_preconditioner
  use module_mpi_comm, only: sync_u, sync_v, apply_boundary_conditions, apply_velocity_bc, apply_vof_bc
  use module_simulation, only: simulation_t
  use module_vtk_writer, only: vtk_writer_t
  implicit none
  include 'mpif.h'

  character(len=256) :: mode, fin, fout
  integer :: ierr, uin, uout

  call MPI_Init(ierr)
  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  ! the sim mode reads its JSON case through the reference's own loader
  if (trim(mode) /= 'sim') open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old', &
       action='read')
  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')

  select case (trim(mode))
  case ('geom')
     call run_geom()
  case ('recon')
     call run_recon()
  case ('advect')
     call run_advect()
  case ('surface')
     call run_surface()
  case ('momentum')
     call run_momentum()
  case ('spstep')
     call run_spstep()
  case ('output')
     call run_output()
  case ('props')
     call run_props()
  case ('jet')
     call run_jet()
  case ('spobs')
     call run_spobs()
  case ('sim')
     call run_sim()
  case ('tadvect')
     call run_tadvect()
  case ('tsp')
     call run_tsp()
  case default
     write(*, '(A,A)') 'unknown mode ', trim(mode)
     call MPI_Abort(MPI_COMM_WORLD, 2, ierr)
  end select

  if (trim(mode) /= 'sim') close(uin)
  close(uout)
  call MPI_Finalize(ierr)

contains

  subroutine run_geom()
    integer :: n, k
    real(dp), allocatable :: xmin(:), xmax(:), ymin(:), ymax(:), gx(:), gy(:), ga(:), gc(:)
    real(dp), allocatable :: area(:), alpha(:)
    read(uin) n
    allocate(xmin(n), xmax(n), ymin(n), ymax(n), gx(n), gy(n), ga(n), gc(n), area(n), alpha(n))
    read(uin) xmin, xmax, ymin, ymax, gx, gy, ga, gc
    do k = 1, n
       area(k) = cut_box_area(xmin(k), xmax(k), ymin(k), ymax(k), gx(k), gy(k), ga(k))
       alpha(k) = compute_alpha(xmin(k), xmax(k), ymin(k), ymax(k), gx(k), gy(k), gc(k))
    end do
    write(uout) area, alpha
  end subroutine run_geom

  subroutine setup(mesh, cfg, nx, ny, lx, ly, bc)
    type(mesh_2d_t), intent(out) :: mesh
    type(config_t), intent(inout) :: cfg
    integer, intent(in) :: nx, ny, bc(4)
    real(dp), intent(in) :: lx, ly
    cfg%nx = nx
    cfg%ny = ny
    cfg%lx = lx
    cfg%ly = ly
    cfg%bc_left%btype = bc(1)
    cfg%bc_right%btype = bc(2)
    cfg%bc_bottom%btype = bc(3)
    cfg%bc_top%btype = bc(4)
    call init_mesh(mesh, MPI_COMM_WORLD, nx, ny, lx, ly, &
                   bc(1) == BC_PERIODIC .or. bc(2) == BC_PERIODIC, &
                   bc(3) == BC_PERIODIC .or. bc(4) == BC_PERIODIC)
  end subroutine setup

  function method_name(m) result(s)
    integer, intent(in) :: m
    character(len=32) :: s
    if (m == 1) then
       s = 'ELVIRA'
    else
       s = 'PLIC'
    end if
  end function method_name

  subroutine run_recon()
    integer :: nx, ny, bc(4), method
    real(dp) :: lx, ly
    type(mesh_2d_t) :: mesh
    type(config_t) :: cfg
    real(dp), allocatable :: C(:, :)
    type(interface_segment_t), allocatable :: seg(:, :)
    read(uin) nx, ny, bc, method
    read(uin) lx, ly
    call setup(mesh, cfg, nx, ny, lx, ly, bc)
    allocate(C(0:nx+1, 0:ny+1), seg(0:nx+1, 0:ny+1))
    read(uin) C
    call reconstruct_interface(method_name(method), mesh, C, seg)
    write(uout) seg%nx, seg%ny, seg%alpha
  end subroutine run_recon

  subroutine run_advect()
    integer :: nx, ny, bc(4), method, two_phase, nsteps, s
    real(dp) :: lx, ly, dt, rho1, rho2, mu1, mu2
    type(mesh_2d_t) :: mesh
    type(config_t) :: cfg
    type(flow_fields_t) :: fields
    read(uin) nx, ny, bc, method, two_phase, nsteps
    read(uin) lx, ly, dt, rho1, rho2, mu1, mu2
    call setup(mesh, cfg, nx, ny, lx, ly, bc)
    cfg%advection_method = 'WEYMOUTH_YUE'
    cfg%reconstruction_method = method_name(method)
    cfg%enable_two_phase = (two_phase /= 0)
    cfg%rho1 = rho1
    cfg%rho2 = rho2
    cfg%mu1 = mu1
    cfg%mu2 = mu2
    cfg%use_markers = .false.
    cfg%enable_svof_smoothing = .false.
    cfg%enable_diffusion_smoothing = .false.
    cfg%enable_surface_tension = .false.
    cfg%compressible = .false.
    call fields%allocate(mesh)
    fields%rho = 1.0_dp
    fields%rho_u = 1.0_dp
    fields%rho_v = 1.0_dp
    fields%mu = 1.0_dp
    fields%mu_node = 1.0_dp
    read(uin) fields%C, fields%u, fields%v
    do s = 1, nsteps
       call advect_vof(fields, mesh, cfg, dt)
       write(uout) fields%C
    end do
    write(uout) fields%rho, fields%rho_u, fields%rho_v, fields%mu, fields%mu_node
  end subroutine run_advect

  ! The advect mode's inputs; nsteps of advect_vof timed by the wall clock
  ! (MPI_Wtime); writes the seconds and the final liquid volume.
  subroutine run_tadvect()
    integer :: nx, ny, bc(4), method, two_phase, nsteps, s
    real(dp) :: lx, ly, dt, rho1, rho2, mu1, mu2, t0, t1
    type(mesh_2d_t) :: mesh
    type(config_t) :: cfg
    type(flow_fields_t) :: fields
    read(uin) nx, ny, bc, method, two_phase, nsteps
    read(uin) lx, ly, dt, rho1, rho2, mu1, mu2
    call setup(mesh, cfg, nx, ny, lx, ly, bc)
    cfg%advection_method = 'WEYMOUTH_YUE'
    cfg%reconstruction_method = method_name(method)
    cfg%enable_two_phase = (two_phase /= 0)
    cfg%rho1 = rho1
    cfg%rho2 = rho2
    cfg%mu1 = mu1
    cfg%mu2 = mu2
    cfg%use_markers = .false.
    cfg%enable_svof_smoothing = .false.
    cfg%enable_diffusion_smoothing = .false.
    cfg%enable_surface_tension = .false.
    cfg%compressible = .false.
    call fields%allocate(mesh)
    fields%rho = 1.0_dp
    fields%rho_u = 1.0_dp
    fields%rho_v = 1.0_dp
    fields%mu = 1.0_dp
    fields%mu_node = 1.0_dp
    read(uin) fields%C, fields%u, fields%v
    t0 = MPI_Wtime()
    do s = 1, nsteps
       call advect_vof(fields, mesh, cfg, dt)
    end do
    t1 = MPI_Wtime()
    write(uout) t1 - t0, sum(fields%C(1:nx, 1:ny))
  end subroutine run_tadvect

  ! One standard-projection step without convection, timed by the wall
  ! clock: the predictor calls alone (on a copy, as spobs replays them), then
  ! solve_standard_projection (predictor + the Poisson block with its
  ! hard-coded CG + SGS to 1e-7 + the correction).  Writes both times and
  ! the iteration counts.
  subroutine run_tsp()
    integer :: nx, ny, bc(4), iters, iters_u
    real(dp) :: lx, ly, dt, rho1, rho2, mu1, mu2, gx, gy, t0, t1, t2
    type(simulation_t) :: sim
    type(flow_fields_t) :: keep
    real(dp), allocatable :: rhs(:), state(:)
    read(uin) nx, ny, bc
    read(uin) lx, ly, dt, rho1, rho2, mu1, mu2, gx, gy
    call setup(sim%mesh, sim%config, nx, ny, lx, ly, bc)
    sim%config%solver_type = SOLVER_STANDARD_PROJECTION
    sim%config%enable_two_phase = .true.
    sim%config%rho1 = rho1
    sim%config%rho2 = rho2
    sim%config%mu1 = mu1
    sim%config%mu2 = mu2
    sim%config%enable_surface_tension = .false.
    sim%config%use_markers = .false.
    sim%config%enable_svof_smoothing = .false.
    sim%config%enable_diffusion_smoothing = .false.
    sim%config%compressible = .false.
    sim%config%time_scheme = 1
    sim%config%enable_convection = .false.
    sim%config%gx = gx
    sim%config%gy = gy
    sim%config%linear_solver = 1
    sim%config%precond_type = 1
    sim%config%use_precond = .true.
    sim%config%tol = 1.0e-10_dp
    sim%config%max_iter = 1000
    sim%config%enable_energy = .false.
    sim%config%enable_ssp_rk3_momentum = .false.
    sim%config%obs_enabled = .false.
    call sim%fields%allocate(sim%mesh)
    call sim%ns_solver%init(sim%mesh, sim%config)
    read(uin) sim%fields%C, sim%fields%u, sim%fields%v
    sim%fields%p = 0.0_dp
    call update_physical_properties(sim%fields, sim%mesh, sim%config)
    sim%fields%u_nm1 = sim%fields%u
    sim%fields%v_nm1 = sim%fields%v
    sim%fields%has_nm1 = .false.
    sim%fields%u_old = sim%fields%u
    sim%fields%v_old = sim%fields%v
    call apply_velocity_bc(sim%fields, sim%mesh, sim%config, .false.)
    keep = sim%fields
    t0 = MPI_Wtime()
    call sim%ns_solver%assemble_preconditioner_blocks(sim%mesh, sim%config, sim%fields, dt)
    allocate(rhs(sim%ns_solver%n_owned_total), state(sim%ns_solver%n_owned_total))
    call sim%ns_solver%assemble_rhs(sim%mesh, sim%config, sim%fields, dt, rhs)
    call sim%ns_solver%pack_vector(sim%mesh, sim%config, sim%fields, state)
    call solve_linear_system(sim%ns_solver, sim%mesh, sim%config, sim%fields, dt, rhs, state, iters_u)
    call sim%ns_solver%unpack_vector(sim%mesh, sim%config, state, sim%fields, .false.)
    t1 = MPI_Wtime()
    sim%fields = keep
    call sim%solve_standard_projection(dt, iters)
    t2 = MPI_Wtime()
    write(uout) t1 - t0, t2 - t1, iters, iters_u
  end subroutine run_tsp

  subroutine run_surface()
    integer :: nx, ny, bc(4), method, svof, diff, dsteps, tension, nsteps, s
    real(dp) :: lx, ly, dt, sigma, width
    type(mesh_2d_t) :: mesh
    type(config_t) :: cfg
    type(flow_fields_t) :: fields
    read(uin) nx, ny, bc, method, svof, diff, dsteps, tension, nsteps
    read(uin) lx, ly, dt, sigma, width
    call setup(mesh, cfg, nx, ny, lx, ly, bc)
    cfg%advection_method = 'WEYMOUTH_YUE'
    cfg%reconstruction_method = method_name(method)
    cfg%enable_two_phase = .false.
    cfg%use_markers = .false.
    cfg%enable_svof_smoothing = (svof /= 0)
    cfg%svof_filter_width = width
    cfg%enable_diffusion_smoothing = (diff /= 0)
    cfg%diffusion_smoothing_steps = dsteps
    cfg%enable_surface_tension = (tension /= 0)
    cfg%sigma = sigma
    cfg%compressible = .false.
    call fields%allocate(mesh)
    fields%rho = 1.0_dp
    fields%rho_u = 1.0_dp
    fields%rho_v = 1.0_dp
    fields%mu = 1.0_dp
    fields%mu_node = 1.0_dp
    read(uin) fields%C, fields%u, fields%v
    do s = 1, nsteps
       call advect_vof(fields, mesh, cfg, dt)
       write(uout) fields%C, fields%C_s, fields%C_tilde, fields%curvature, fields%fx_sigma, fields%fy_sigma
    end do
  end subroutine run_surface

  subroutine run_momentum()
    integer :: nx, ny, bc(4), tscheme, has_nm1, conv, scheme, lin, solver, precond, maxit, iters, stype
    real(dp) :: lx, ly, dt, gx, gy, tol, bcu(4), bcv(4)
    type(mesh_2d_t) :: mesh
    type(config_t) :: cfg
    type(flow_fields_t) :: fields
    type(navier_stokes_t) :: ns
    real(dp), allocatable :: rhs(:), x1(:), x2(:), y1(:), y2(:), z1(:), state(:)
    read(uin) nx, ny, bc, tscheme, has_nm1, conv, scheme, lin, solver, precond, maxit, stype
    read(uin) lx, ly, dt, gx, gy, tol, bcu, bcv
    call setup(mesh, cfg, nx, ny, lx, ly, bc)
    cfg%solver_type = stype
    cfg%time_scheme = tscheme
    cfg%enable_convection = (conv /= 0)
    cfg%convective_scheme = scheme
    cfg%convective_linearization = lin
    cfg%gx = gx
    cfg%gy = gy
    cfg%linear_solver = solver
    cfg%precond_type = precond
    cfg%use_precond = .true.
    cfg%tol = tol
    cfg%max_iter = maxit
    cfg%obs_enabled = .false.
    cfg%enable_energy = .false.
    cfg%enable_ssp_rk3_momentum = .false.
    cfg%compressible = .false.
    cfg%bc_left%u = bcu(1)
    cfg%bc_right%u = bcu(2)
    cfg%bc_bottom%u = bcu(3)
    cfg%bc_top%u = bcu(4)
    cfg%bc_left%v = bcv(1)
    cfg%bc_right%v = bcv(2)
    cfg%bc_bottom%v = bcv(3)
    cfg%bc_top%v = bcv(4)
    call fields%allocate(mesh)
    call ns%init(mesh, cfg)
    read(uin) fields%u, fields%v, fields%u_old, fields%v_old, fields%u_nm1, fields%v_nm1
    read(uin) fields%rho_u, fields%rho_v, fields%mu, fields%mu_node, fields%fx_sigma, fields%fy_sigma, fields%p
    fields%has_nm1 = (has_nm1 /= 0)
    allocate(rhs(ns%n_owned_total), x1(ns%n_owned_total), x2(ns%n_owned_total), z1(ns%n_owned_total))
    allocate(y1(ns%n_owned_total), y2(ns%n_owned_total), state(ns%n_owned_total))
    read(uin) x1, x2
    call ns%assemble_preconditioner_blocks(mesh, cfg, fields, dt)
    if (stype == SOLVER_MONOLITHIC) then
       write(uout) ns%diag_u, ns%diag_v, ns%diag_p
    else
       write(uout) ns%diag_u, ns%diag_v
    end if
    call ns%assemble_rhs(mesh, cfg, fields, dt, rhs)
    write(uout) rhs
    call ns%apply_system_matrix(mesh, cfg, fields, dt, x1, y1)
    call ns%apply_system_matrix(mesh, cfg, fields, dt, x2, y2)
    write(uout) y1, y2
    call apply_preconditioner(ns, mesh, cfg, x1, z1)
    write(uout) z1
    call ns%pack_vector(mesh, cfg, fields, state)
    call solve_linear_system(ns, mesh, cfg, fields, dt, rhs, state, iters)
    call ns%unpack_vector(mesh, cfg, state, fields, .false.)
    if (stype /= SOLVER_MONOLITHIC) then
       call sync_u(mesh, fields%u)
       call sync_v(mesh, fields%v)
       call apply_boundary_conditions(mesh, cfg, fields, .false.)
    end if
    write(uout) fields%u, fields%v, fields%p
    write(uout) iters
  end subroutine run_momentum

  subroutine run_spstep()
    integer :: nx, ny, bc(4), method, tscheme, scheme, lin, solver, precond, maxit, nsteps, s, iters
    real(dp) :: lx, ly, dt, rho1, rho2, mu1, mu2, sigma, gx, gy, tol, bcu(4), bcv(4)
    type(simulation_t) :: sim
    read(uin) nx, ny, bc, method, tscheme, scheme, lin, solver, precond, maxit, nsteps
    read(uin) lx, ly, dt, rho1, rho2, mu1, mu2, sigma, gx, gy, tol, bcu, bcv
    call setup(sim%mesh, sim%config, nx, ny, lx, ly, bc)
    sim%config%solver_type = SOLVER_STANDARD_PROJECTION
    sim%config%advection_method = 'WEYMOUTH_YUE'
    sim%config%reconstruction_method = method_name(method)
    sim%config%enable_two_phase = .true.
    sim%config%rho1 = rho1
    sim%config%rho2 = rho2
    sim%config%mu1 = mu1
    sim%config%mu2 = mu2
    sim%config%sigma = sigma
    sim%config%enable_surface_tension = (sigma > 0.0_dp)
    sim%config%use_markers = .false.
    sim%config%enable_svof_smoothing = .false.
    sim%config%enable_diffusion_smoothing = .false.
    sim%config%compressible = .false.
    sim%config%time_scheme = tscheme
    sim%config%enable_convection = .true.
    sim%config%convective_scheme = scheme
    sim%config%convective_linearization = lin
    sim%config%gx = gx
    sim%config%gy = gy
    sim%config%linear_solver = solver
    sim%config%precond_type = precond
    sim%config%use_precond = .true.
    sim%config%tol = tol
    sim%config%max_iter = maxit
    sim%config%obs_enabled = .false.
    sim%config%enable_energy = .false.
    sim%config%enable_ssp_rk3_momentum = .false.
    sim%config%bc_left%u = bcu(1)
    sim%config%bc_right%u = bcu(2)
    sim%config%bc_bottom%u = bcu(3)
    sim%config%bc_top%u = bcu(4)
    sim%config%bc_left%v = bcv(1)
    sim%config%bc_right%v = bcv(2)
    sim%config%bc_bottom%v = bcv(3)
    sim%config%bc_top%v = bcv(4)
    call sim%fields%allocate(sim%mesh)
    call sim%ns_solver%init(sim%mesh, sim%config)
    read(uin) sim%fields%C, sim%fields%u, sim%fields%v
    sim%fields%p = 0.0_dp
    call update_physical_properties(sim%fields, sim%mesh, sim%config)
    do s = 1, nsteps
       if (s > 1) then
          sim%fields%u_nm1 = sim%fields%u_old
          sim%fields%v_nm1 = sim%fields%v_old
          sim%fields%has_nm1 = .true.
       else
          sim%fields%u_nm1 = sim%fields%u
          sim%fields%v_nm1 = sim%fields%v
          sim%fields%has_nm1 = .false.
       end if
       sim%fields%u_old = sim%fields%u
       sim%fields%v_old = sim%fields%v
       sim%fields%p_old = sim%fields%p
       call advect_vof(sim%fields, sim%mesh, sim%config, dt)
       call apply_velocity_bc(sim%fields, sim%mesh, sim%config, .false.)
       call apply_vof_bc(sim%fields, sim%mesh, sim%config)
       call sim%solve_standard_projection(dt, iters)
       write(uout) sim%fields%C, sim%fields%u, sim%fields%v, sim%fields%p
       write(uout) iters
    end do
  end subroutine run_spstep

  subroutine run_props()
    integer :: nx, ny, bc(4), svof
    real(dp) :: lx, ly, rho1, rho2, mu1, mu2
    type(mesh_2d_t) :: mesh
    type(config_t) :: cfg
    type(flow_fields_t) :: fields
    read(uin) nx, ny, bc, svof
    read(uin) lx, ly, rho1, rho2, mu1, mu2
    call setup(mesh, cfg, nx, ny, lx, ly, bc)
    cfg%enable_two_phase = .true.
    cfg%enable_svof_smoothing = (svof /= 0)
    cfg%compressible = .false.
    cfg%rho1 = rho1
    cfg%rho2 = rho2
    cfg%mu1 = mu1
    cfg%mu2 = mu2
    call fields%allocate(mesh)
    read(uin) fields%C, fields%C_s
    call update_physical_properties(fields, mesh, cfg)
    write(uout) fields%C, fields%rho, fields%rho_u, fields%rho_v, fields%mu, fields%mu_node
  end subroutine run_props

  subroutine run_jet()
    integer :: nx, ny, bc(4), method, nsteps, s
    real(dp) :: lx, ly, dt, jc, jw
    type(mesh_2d_t) :: mesh
    type(config_t) :: cfg
    type(flow_fields_t) :: fields
    read(uin) nx, ny, bc, method, nsteps
    read(uin) lx, ly, dt, jc, jw
    call setup(mesh, cfg, nx, ny, lx, ly, bc)
    cfg%advection_method = 'WEYMOUTH_YUE'
    cfg%reconstruction_method = method_name(method)
    cfg%enable_two_phase = .false.
    cfg%use_markers = .false.
    cfg%enable_svof_smoothing = .false.
    cfg%enable_diffusion_smoothing = .false.
    cfg%enable_surface_tension = .false.
    cfg%compressible = .false.
    cfg%bc_top%profile = 'JET_INLET'
    cfg%bc_top%jet_center = jc
    cfg%bc_top%jet_width = jw
    call fields%allocate(mesh)
    read(uin) fields%C, fields%u, fields%v
    do s = 1, nsteps
       call advect_vof(fields, mesh, cfg, dt)
       write(uout) fields%C
    end do
  end subroutine run_jet

  subroutine run_spobs()
    integer :: nx, ny, bc(4), obs_type, iters, iters_u
    real(dp) :: lx, ly, dt, rho1, rho2, mu1, mu2, gx, gy, ob(7)
    type(simulation_t) :: sim
    type(flow_fields_t) :: keep
    real(dp), allocatable :: rhs(:), state(:)
    read(uin) nx, ny, bc, obs_type
    read(uin) lx, ly, dt, rho1, rho2, mu1, mu2, gx, gy, ob
    call setup(sim%mesh, sim%config, nx, ny, lx, ly, bc)
    sim%config%solver_type = SOLVER_STANDARD_PROJECTION
    sim%config%enable_two_phase = .true.
    sim%config%rho1 = rho1
    sim%config%rho2 = rho2
    sim%config%mu1 = mu1
    sim%config%mu2 = mu2
    sim%config%enable_surface_tension = .false.
    sim%config%use_markers = .false.
    sim%config%enable_svof_smoothing = .false.
    sim%config%enable_diffusion_smoothing = .false.
    sim%config%compressible = .false.
    sim%config%time_scheme = 1
    sim%config%enable_convection = .false.
    sim%config%gx = gx
    sim%config%gy = gy
    sim%config%linear_solver = 1
    sim%config%precond_type = 1
    sim%config%use_precond = .true.
    sim%config%tol = 1.0e-12_dp
    sim%config%max_iter = 1000
    sim%config%enable_energy = .false.
    sim%config%enable_ssp_rk3_momentum = .false.
    sim%config%obs_enabled = .true.
    sim%config%obs_type = obs_type
    sim%config%obs_x0 = ob(1)
    sim%config%obs_y0 = ob(2)
    sim%config%obs_x1 = ob(3)
    sim%config%obs_y1 = ob(4)
    sim%config%obs_xc = ob(5)
    sim%config%obs_yc = ob(6)
    sim%config%obs_radius = ob(7)
    sim%config%obs_perm = 0.0_dp
    call sim%fields%allocate(sim%mesh)
    call sim%ns_solver%init(sim%mesh, sim%config)
    read(uin) sim%fields%C, sim%fields%u, sim%fields%v
    sim%fields%p = 0.0_dp
    call update_physical_properties(sim%fields, sim%mesh, sim%config)
    sim%fields%u_nm1 = sim%fields%u
    sim%fields%v_nm1 = sim%fields%v
    sim%fields%has_nm1 = .false.
    sim%fields%u_old = sim%fields%u
    sim%fields%v_old = sim%fields%v
    call apply_velocity_bc(sim%fields, sim%mesh, sim%config, .false.)
    ! the predictor calls of solve_standard_projection on a copy: u*, v*
    keep = sim%fields
    call sim%ns_solver%assemble_preconditioner_blocks(sim%mesh, sim%config, sim%fields, dt)
    allocate(rhs(sim%ns_solver%n_owned_total), state(sim%ns_solver%n_owned_total))
    call sim%ns_solver%assemble_rhs(sim%mesh, sim%config, sim%fields, dt, rhs)
    call sim%ns_solver%pack_vector(sim%mesh, sim%config, sim%fields, state)
    call solve_linear_system(sim%ns_solver, sim%mesh, sim%config, sim%fields, dt, rhs, state, iters_u)
    call sim%ns_solver%unpack_vector(sim%mesh, sim%config, state, sim%fields, .false.)
    call sync_u(sim%mesh, sim%fields%u)
    call sync_v(sim%mesh, sim%fields%v)
    call apply_boundary_conditions(sim%mesh, sim%config, sim%fields, .false.)
    write(uout) sim%fields%u, sim%fields%v, sim%fields%rho_u, sim%fields%rho_v
    sim%fields = keep
    call sim%solve_standard_projection(dt, iters)
    write(uout) sim%fields%p, sim%fields%u, sim%fields%v
    write(uout) iters, iters_u
  end subroutine run_spobs

  subroutine run_sim()
    type(simulation_t) :: sim
    if (.not. sim%init(MPI_COMM_WORLD, trim(fin))) then
       write(*, '(A)') 'sim%init failed'
       call MPI_Abort(MPI_COMM_WORLD, 3, ierr)
    end if
    call sim%run()
    write(uout) 1
  end subroutine run_sim

  subroutine run_output()
    integer :: nx, ny, bc(4), nsnap, s, step
    real(dp) :: lx, ly, t, dt
    type(mesh_2d_t) :: mesh
    type(config_t) :: cfg
    type(flow_fields_t) :: fields
    type(vtk_writer_t) :: vtk
    read(uin) nx, ny, bc, nsnap
    read(uin) lx, ly
    call setup(mesh, cfg, nx, ny, lx, ly, bc)
    call fields%allocate(mesh)
    call vtk%init(mesh, trim(fout) // '_dir', 'case')
    do s = 1, nsnap
       read(uin) step
       read(uin) t, dt
       read(uin) fields%C, fields%u, fields%v, fields%p, fields%rho, fields%mu, fields%T, fields%curvature, &
            fields%C_s
       call vtk%write_step(mesh, fields, step, t)
       call vtk%write_diagnostics(mesh, fields, step, t, dt)
    end do
    call vtk%finalize()
    write(uout) nsnap
  end subroutine run_output

end program covof_ref_harness
