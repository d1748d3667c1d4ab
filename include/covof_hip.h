/*
 * covof_hip.h -- C-ABI of the MI355X (gfx950) hot path of COVOF 2D.
 *
 * Drop-in replacements for the per-timestep hot path of the reference Fortran
 * solver: Weymouth-Yue VOF advection with PLIC/ELVIRA reconstruction and the
 * steps that follow it in advect_vof, the SP velocity predictor (momentum
 * block), the variable-density pressure-Poisson solve of the standard
 * projection, and the velocity correction.  Plain C types only (pointers, ints, doubles), so the
 * Fortran driver binds them with ISO_C_BINDING and Python with ctypes.
 *
 * Conventions
 *  - Field layout is the reference's: Fortran F(0:nx+1, 0:ny+1), column-major,
 *    i fastest (src/module_fields.f90:129-169): element (i,j) is at offset
 *    j*(nx+2)+i.  u(i,j) lives on the face x_face(i), v(i,j) on y_face(j)
 *    (src/module_fields.f90:33-35).  Poisson work arrays are (1:nx,1:ny):
 *    offset (j-1)*nx+(i-1).
 *  - Boundary types, linear-solver and preconditioner selectors are the
 *    reference's integers (src/module_constants.f90:25-31, :53-77).
 *  - Entry points ending in _d take DEVICE pointers and run on the context's
 *    HIP stream (set with covof_ctx_set_stream); they return once the work is
 *    queued unless they must read a result back (iteration counts).  Entry
 *    points without the suffix take HOST arrays (the Fortran drop-in), copy in,
 *    run, copy out and synchronise.
 *  - Every call returns an int status (COVOF_OK on success) and never aborts.
 *  - One context per thread/stream; calls on one context are serialised.
 */
#ifndef COVOF_HIP_H
#define COVOF_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (the reference exits silently; here every outcome is explicit). */
#define COVOF_OK 0
#define COVOF_NOT_CONVERGED 1 /* max_iter reached */
#define COVOF_BREAKDOWN 2     /* |d.Ad| or |r.z| (or BiCGSTAB(2) scalars) underflowed */
#define COVOF_NONFINITE 3     /* NaN/Inf in the residual */
#define COVOF_BAD_ARGUMENT 4
#define COVOF_HIP_ERROR 5

/* Reference selector values (src/module_constants.f90:25-31, :53-77). */
#define COVOF_BC_DIRICHLET 1
#define COVOF_BC_NEUMANN 2
#define COVOF_BC_PERIODIC 3
#define COVOF_BC_FREESLIP 4
#define COVOF_LINEAR_BICGSTAB2 1
#define COVOF_LINEAR_CG 3
#define COVOF_PRECOND_NONE 0
#define COVOF_PRECOND_JACOBI 1
#define COVOF_PRECOND_SSOR 2 /* symmetric Gauss-Seidel, lexicographic (reference SP) */
#define COVOF_PRECOND_BLOCK_TRIANGULAR 4 /* Schur-SSOR then velocity SSOR (monolithic) */
#define COVOF_SOLVER_MONOLITHIC 2
/* Curvature of the CSF forces (covof_curvature_csf_method_d). */
#define COVOF_CURVATURE_NORMALS 0
#define COVOF_CURVATURE_HEIGHT_FUNCTION 1
#define COVOF_SOLVER_STANDARD_PROJECTION 3
/* Parallel preconditioners with no reference counterpart (values clear of
 * the reference's 0-5; they take the place PRECOND_CUSTOM = 5 gives a user
 * hook, src/module_linear_solvers.f90:22-41): */
#define COVOF_PRECOND_SGS_RB 16 /* one symmetric red-black Gauss-Seidel sweep */
#define COVOF_PRECOND_MG 17     /* one geometric multigrid V-cycle (RB-GS smoothing) */
/* Arithmetic of the Poisson solve. */
#define COVOF_PRECISION_FP64 0
#define COVOF_PRECISION_MIXED 1 /* FP32 inner Krylov, FP64 iterative refinement */

typedef struct covof_ctx covof_ctx;

/* Mesh + device buffers.  Replaces init_mesh (src/module_mesh.f90:67-144) for
 * one rank and the per-call allocations of advect_vof_weymouth_yue
 * (src/module_interface_vof.f90:1182-1186) and of the SP Poisson block
 * (src/module_simulation.f90:2477-2479): all buffers are made here, once.
 * bc[4] = left, right, bottom, top. */
int covof_ctx_create(int nx, int ny, double lx, double ly, const int *bc, covof_ctx **out);
int covof_ctx_destroy(covof_ctx *ctx);
/* Use an existing HIP stream (hipStream_t passed as void*); NULL = own stream. */
int covof_ctx_set_stream(covof_ctx *ctx, void *stream);
int covof_ctx_synchronize(covof_ctx *ctx);
/* Message of the last error seen on this thread (static storage). */
const char *covof_last_error(void);
/* Summation order of the FP64 Krylov inner products and of the pressure mean.
 * COVOF_REDUCTION_TREE (the default): deterministic parallel tree sums (the
 * same inputs give the same bits run to run).  COVOF_REDUCTION_REFERENCE: the
 * reference's order -- sum(x*y) / dot_product over the column-major arrays,
 * one accumulator in index order (global_dot_product_2d and the p mean,
 * src/module_simulation.f90:2577-2585, :2760-2774; global_dot_product,
 * src/module_mpi_comm.f90:820-833, as the Krylov loops of
 * src/module_linear_solvers.f90 call it) -- so that FP64 solves reproduce the
 * reference's iterates bit for bit; each sum is then one sequential chain on
 * the GPU (about 0.1 ms per 32k terms).  The FP32 recurrence of the mixed
 * mode always uses tree sums. */
#define COVOF_REDUCTION_TREE 0
#define COVOF_REDUCTION_REFERENCE 1
int covof_ctx_set_reduction_order(covof_ctx *ctx, int order);

/* Tuning of the context's Poisson solves (no reference counterpart).
 * Setting one drops the context's captured solver cycles and multigrid
 * coarse operators (made again by the next solve).  value: */
#define COVOF_OPT_MIXED_STALL_WINDOW 1 /* reliable updates without a 10 % gain before the mixed mode continues
                                          in FP64; 0 (default): 48 (Jacobi / none), 2 (MG) */
#define COVOF_OPT_MIXED_CYCLE 2        /* FP32 iterations per reliable update, even, >= 2; 0 (default): 128
                                          (Jacobi / none), 32 (SGS_RB), 8 (MG) */
#define COVOF_OPT_MIXED_LOOP 3         /* 0 (default): two-kernel PCG; 1: the single-reduction
                                          Chronopoulos-Gear loop (mixed Jacobi only) */
#define COVOF_OPT_MG_COARSE 4          /* coarse faces of PRECOND_MG: 0 (default) averaged, 1 in series */
#define COVOF_OPT_MG_SWEEPS 5          /* red-black sweeps before and after the coarse correction, 1..4
                                          (default 1) */
#define COVOF_OPT_MG_OMEGA 6           /* weight of the coarse-grid correction, (0, 2] (default 1.8) */
#define COVOF_OPT_MG_FUSED 7           /* 1 (default): with one sweep, each multi-workgroup level of the V-cycle
                                          in two passes (smoothing + restriction, prolongation + smoothing);
                                          0: five launches per level.  Same values either way. */
#define COVOF_OPT_MIXED_MG_FORM 8      /* PRECISION_MIXED with PRECOND_MG: 0 (default) FP64 CG whose V-cycle
                                          runs in FP32; 1 the FP32 Krylov recurrence with FP64 reliable
                                          updates (the form of the other preconditioners) */
#define COVOF_OPT_MG_PROJECT 9         /* 1 (default): with a singular operator (no BC_NEUMANN side) the V-cycle's
                                          coarsest solve is mean-free on both sides; 0: plain sweeps */
int covof_ctx_set_solver_option(covof_ctx *ctx, int option, double value);
/* Makes ahead the work arrays a path otherwise makes on its first use, so
 * that the calls of a time step allocate nothing (no reference counterpart;
 * the reference allocates its temporaries every call).  parts: an OR of */
#define COVOF_RESERVE_MOMENTUM 1      /* the momentum / monolithic block (vectors, SSOR and Schur blocks) */
#define COVOF_RESERVE_MULTIGRID 2     /* the PRECOND_MG / SGS_RB hierarchy buffers, FP64 and FP32 */
#define COVOF_RESERVE_SGS 4           /* the exact SGS (PRECOND_SSOR) strip records */
#define COVOF_RESERVE_MIXED_LOOP 8    /* the single-reduction mixed loop's vectors */
#define COVOF_RESERVE_SURFACE 16      /* curvature / smoothing work arrays and the SVOF strip records */
#define COVOF_RESERVE_ADJOINT 32      /* the Weymouth-Yue step adjoint's work arrays */
#define COVOF_RESERVE_HOST_MIRRORS 64 /* device copies behind the host-array entry points */
#define COVOF_RESERVE_ALL 127
int covof_ctx_reserve(covof_ctx *ctx, int parts);
/* The split of the context's last mixed-precision CG solve: iterations of
 * the FP32 recurrence and of the FP64 continuation that follows a stall (0
 * when it did not run); for the FP64 CG with an FP32 V-cycle (MG's default
 * mixed form) 0 and its iterations.  After a continuation the solve's relres is the
 * recomputed true FP64 residual and its status the FP64 PCG's (stopped on
 * its recurrence residual), so where the system's attainable accuracy is
 * above tol the relres can exceed tol with status COVOF_OK. */
int covof_ctx_mixed_solve_info(covof_ctx *ctx, int *fp32_iters, int *fp64_iters);

/* A solid obstacle: config_t's obs_enabled, obs_type (1 box [x0,x1]x[y0,y1],
 * 2 disc of centre (xc, yc) and radius r), obs_perm (src/module_config.f90:
 * 121-129).  With obs_perm <= 1e-12 every velocity face whose point
 * (x_face, y_center) / (x_center, y_face) lies in the obstacle
 * (is_in_obstacle, src/module_mesh.f90:203-222; is_u/v_face_in_obstacle,
 * src/module_navier_stokes.f90:82-131) is a wall of the SP Poisson assembly
 * and is skipped by the velocity correction (src/module_simulation.f90:
 * 2491-2528, :2596-2632). */
typedef struct covof_obstacle {
  int enabled, type;
  double x0, y0, x1, y1, xc, yc, r, perm;
} covof_obstacle;
/* The context's obstacle (NULL: none, the default).  The momentum block does
 * not support obstacles (covof_momentum_set_fields_d refuses while one is set). */
int covof_ctx_set_obstacle(covof_ctx *ctx, const covof_obstacle *obs);
/* The top side's VOF boundary profile of apply_vof_bc (cfg%bc_top%profile,
 * src/module_mpi_comm.f90:655-677): "UNIFORM" (default: zero-gradient copy)
 * or "JET_INLET": the top halo row holds C = 0 where |x_center - xc| <= w
 * and 1 elsewhere, xc = jet_center (<= 0: lx/2), w = jet_width/2 (jet_width
 * <= 0: 0.04); it applies when the y direction is not periodic.  The VOF
 * transport writes this halo after each pass. */
int covof_ctx_set_vof_top_profile(covof_ctx *ctx, const char *profile, double jet_center, double jet_width);

/* --- interface geometry (src/module_interface_reconstruction.f90) ----------- */

/* Batched cut_box_area (:33-101) and compute_alpha (:105-147) on the device:
 * area[k] = cut_box_area(xmin..ymax, nx, ny, alpha), alpha_out[k] =
 * compute_alpha(xmin..ymax, nx, ny, c).  Device arrays of length n. */
int covof_geom_batch_d(covof_ctx *ctx, int n, const double *xmin, const double *xmax,
                       const double *ymin, const double *ymax, const double *gnx,
                       const double *gny, const double *alpha, const double *c,
                       double *area, double *alpha_out);

/* reconstruct_interface(method_name, mesh, C, segments)
 * (src/module_interface_reconstruction.f90:580-683).  method "PLIC" (default,
 * any other name) or "ELVIRA".  Outputs the (nx, ny, alpha) of every cell of
 * the halo grid, as the reference's interface_segment_t array (:23-27), split
 * into three halo-layout arrays. */
int covof_reconstruct_interface_d(covof_ctx *ctx, const char *method, const double *C,
                                  double *seg_nx, double *seg_ny, double *seg_alpha);

/* Allocation-free variant of covof_reconstruct_interface_d, with no
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  It needs no
 * context: the mesh comes as covof_ctx_create's arguments, the work is queued
 * on `stream` (a hipStream_t, NULL = the null stream) of the current device,
 * and it returns once queued.  Its only scratch (the per-tile lists of mixed
 * cells) is a static 4.2 MB device array, one per GPU, held until the process
 * ends; the segments are written straight into seg_nx, seg_ny, seg_alpha
 * (distinct from C and from each other).  The mesh may span at most 1024
 * tiles of 64 x 16 cells (ceil(nx/64) * ceil(ny/16) <= 1024: 1024 x 1024,
 * 2048 x 512, ...); larger ones return COVOF_BAD_ARGUMENT.  Calls that
 * overlap on one GPU share the scratch and race: serialise them. */
int covof_reconstruct_interface_noalloc_d(int nx, int ny, double lx, double ly, const int *bc, const char *method,
                                          const double *C, double *seg_nx, double *seg_ny, double *seg_alpha,
                                          void *stream);

/* --- VOF transport (src/module_interface_vof.f90) ---------------------------- */

/* advect_vof(fields, mesh, cfg, dt) (:792-825) with advection_method
 * "WEYMOUTH_YUE" (:1162-1377), nsteps times.  C (halo layout) is advanced in
 * place; u, v are the face velocities (halo layout), constant over the call.
 * *step_count is the reference's Strang counter vof_step_count (:27): the
 * x-pass goes first when it is even; it is incremented once per step.
 * If two_phase != 0, update_physical_properties (:479-616) follows every step
 * (physical-corner fix of C, then rho, rho_u, rho_v, mu, mu_node, all halo
 * layout; they may be NULL when two_phase == 0).  Other advection methods
 * return COVOF_BAD_ARGUMENT. */
int covof_vof_advect_d(covof_ctx *ctx, double *C, const double *u, const double *v, double dt,
                       const char *advection_method, const char *reconstruction_method,
                       int *step_count, int nsteps, int two_phase, double rho1, double rho2,
                       double mu1, double mu2, double *rho, double *rho_u, double *rho_v,
                       double *mu, double *mu_node);
/* Host-array variant of covof_vof_advect_d (the Fortran drop-in for advect_vof). */
int covof_vof_advect(covof_ctx *ctx, double *C, const double *u, const double *v, double dt,
                     const char *advection_method, const char *reconstruction_method,
                     int *step_count, int nsteps, int two_phase, double rho1, double rho2,
                     double mu1, double mu2, double *rho, double *rho_u, double *rho_v,
                     double *mu, double *mu_node);
/* update_physical_properties (src/module_interface_vof.f90:479-616),
 * two-phase incompressible branch. */
int covof_update_properties_d(covof_ctx *ctx, double *C, double rho1, double rho2, double mu1,
                              double mu2, double *rho, double *rho_u, double *rho_v, double *mu,
                              double *mu_node);

/* update_physical_properties with the SVOF branch (src/module_interface_vof.f90:
 * 479-616, :540-545, :596-600): C_s non-NULL (cfg%enable_svof_smoothing with
 * the smoothed field of covof_svof_smooth_d) takes rho and mu from the clamped
 * C_s and mu_node from the clamped node mean of C_s, harmonically blended;
 * C_s NULL is covof_update_properties_d. */
int covof_update_properties_cs_d(covof_ctx *ctx, double *C, const double *C_s, double rho1, double rho2,
                                 double mu1, double mu2, double *rho, double *rho_u, double *rho_v, double *mu,
                                 double *mu_node);

/* --- the steps advect_vof runs after the transport (single rank, no markers);
 * halo fields (0:nx+1, 0:ny+1); an output never aliases an input. */

/* compute_curvature_and_csf (src/module_interface_vof.f90:627-789): unit
 * normals of grad G (G = C_tilde with diffusion smoothing, C_s with SVOF
 * smoothing, else C; NULL means C), curvature = -div(n) with its exchanged
 * halo, and the CSF forces fx_sigma on the x-faces (1..nx+1, 1..ny) and
 * fy_sigma on the y-faces (1..nx, 1..ny+1) from the face curvature and the
 * face gradient of C; their other entries are left as they are.  enable == 0
 * (cfg%enable_surface_tension false) or sigma <= 0 zeroes all three arrays. */
int covof_curvature_csf_d(covof_ctx *ctx, const double *G, const double *C, int enable, double sigma,
                          double *curvature, double *fx_sigma, double *fy_sigma);
/* The same with a curvature method: COVOF_CURVATURE_NORMALS (the reference's
 * -div(n), as covof_curvature_csf_d) or COVOF_CURVATURE_HEIGHT_FUNCTION:
 * curvature from 3 x 7 column heights of C (Cummins, Francois & Kothe,
 * Computers & Structures 83 (2005) 425-434) wherever the columns run from
 * empty to full along the normal, the reference's value elsewhere.  The
 * height-function option has no counterpart in the reference (a new feature
 * for the full coupled step's CSF forces); the force assembly is the
 * reference's. */
int covof_curvature_csf_method_d(covof_ctx *ctx, const double *G, const double *C, int enable, double sigma,
                                 int method, double *curvature, double *fx_sigma, double *fy_sigma);
int covof_curvature_csf(covof_ctx *ctx, const double *G, const double *C, int enable, double sigma,
                        double *curvature, double *fx_sigma, double *fy_sigma);
/* vof_smooth_diffusion (:1416-1461): steps explicit diffusion steps of C
 * (steps < 1 means 3, cfg%diffusion_smoothing_steps) into C_tilde, with the
 * reference's halo fill (exchange, zero-gradient copies on physical sides). */
int covof_smooth_diffusion_d(covof_ctx *ctx, const double *C, int steps, double *C_tilde);
int covof_smooth_diffusion(covof_ctx *ctx, const double *C, int steps, double *C_tilde);
/* smooth_svof_helmholtz (:387-476): 25 symmetric Gauss-Seidel iterations of
 * (I - D Lap) C_s = C, D = filter_width * dx * dy (filter_width <= 0 means
 * 1, cfg%svof_filter_width), then the clamp to [0, 1] and the halo.  The
 * recurrence is sequential: it runs in one workgroup (the parity form). */
int covof_svof_smooth_d(covof_ctx *ctx, const double *C, double filter_width, double *C_s);
int covof_svof_smooth(covof_ctx *ctx, const double *C, double filter_width, double *C_s);

/* --- pressure Poisson of the standard projection (src/module_simulation.f90) */

/* Assembly of the variable-density 5-point operator and of b = -div(u*)
 * (:2477-2533) into the context; b (nx*ny, device) may be NULL. */
int covof_poisson_assemble_d(covof_ctx *ctx, const double *u, const double *v,
                             const double *rho_u, const double *rho_v, double dt, double *b);
/* y = A x, the reference's poisson_matvec (:2668-2707); (nx,ny) arrays. */
int covof_poisson_matvec_d(covof_ctx *ctx, const double *x, double *y);
/* z = M^-1 r: NONE, JACOBI (src/module_linear_solvers.f90:88-135) or SSOR =
 * the symmetric Gauss-Seidel poisson_ssor (src/module_simulation.f90:2709-2758);
 * SGS_RB and MG as above (FP64; they need even sizes in periodic directions). */
int covof_poisson_precond_d(covof_ctx *ctx, int precond, const double *r, double *z);
/* Krylov solve of A x = b on the assembled operator, x (in: initial guess,
 * out: solution), (nx,ny) device arrays.
 *   lin_solver CG: the SP PCG loop (:2535-2575): stop when ||r||/||b|| <= tol,
 *     ||b|| < 1e-30 -> 1, exits on |d.Ad| < 1e-35 or |r.z| < 1e-35.
 *   lin_solver BICGSTAB2: src/module_linear_solvers.f90:361-579 (left
 *     preconditioned, stop on the preconditioned residual ratio or 1e-12).
 *   precision MIXED: the Krylov loop runs in FP32 inside an FP64
 *     iterative refinement that recomputes r = b - A x in FP64 and stops on
 *     the FP64 ||r||/||b|| <= tol (CG with NONE, JACOBI, SGS_RB or MG);
 *     when that residual stagnates above tol (no 10 % gain over 48
 *     refinements, 2 of 8 iterations with MG) FP64 PCG continues from the refined x.
 *     With MG the default form is instead FP64 CG whose preconditioner, the
 *     V-cycle, runs in FP32 (COVOF_OPT_MIXED_MG_FORM); its relres is the
 *     true FP64 residual of the returned x.
 * *iters: Krylov iterations (inner iterations summed for MIXED, plus the
 * FP64 continuation's); *relres:
 * final relative residual as defined by the stop rule. */
int covof_poisson_solve_d(covof_ctx *ctx, const double *b, double *x, int lin_solver, int precond,
                          double tol, int maxit, int precision, int *iters, double *relres);
/* The whole Poisson block of sim_solve_standard_projection (:2477-2588):
 * assembly from u*, v*, rho_u, rho_v; x0 = p(1:nx,1:ny); solve; p = x - mean(x)
 * written back to p(1:nx,1:ny), halo wrapped for periodic directions.  The
 * reference hard-codes CG + SSOR, tol 1e-7, max_iter 1000 (:2437-2438). */
int covof_sp_poisson_d(covof_ctx *ctx, const double *u, const double *v, const double *rho_u,
                       const double *rho_v, double *p, double dt, int lin_solver, int precond,
                       double tol, int maxit, int precision, int *iters, double *relres);
int covof_sp_poisson(covof_ctx *ctx, const double *u, const double *v, const double *rho_u,
                     const double *rho_v, double *p, double dt, int lin_solver, int precond,
                     double tol, int maxit, int precision, int *iters, double *relres);
/* Velocity correction u -= (dt/rho_u) dp/dx, v -= (dt/rho_v) dp/dy on
 * non-wall faces (:2590-2654). */
int covof_sp_correct_d(covof_ctx *ctx, const double *p, const double *rho_u, const double *rho_v,
                       double dt, double *u, double *v);
int covof_sp_correct(covof_ctx *ctx, const double *p, const double *rho_u, const double *rho_v,
                     double dt, double *u, double *v);

/* --- SP velocity predictor (the momentum block of the standard projection) - */

/* Options of the momentum block (the config_t fields it reads; reference
 * integers of src/module_constants.f90:81-85, :132-140).  bc_u/bc_v: the
 * Dirichlet values cfg%bc_<side>%u / %v of left, right, bottom, top. */
typedef struct covof_momentum_cfg {
  int time_scheme;              /* 1 EULER, 2 GEAR (BDF2) */
  int has_nm1;                  /* fields%has_nm1: u_nm1, v_nm1 hold the n-1 state */
  int enable_convection;        /* cfg%enable_convection */
  int convective_scheme;        /* 1 UPWIND, 2 CENTERED, 3 TVD_SUPERBEE */
  int convective_linearization; /* 1 PICARD, 2 ADAMS_BASHFORTH, 3 NEWTON */
  double gx, gy;
  double bc_u[4], bc_v[4];
  int solver_type;              /* cfg%solver_type: 0 (unset) or 3 STANDARD_PROJECTION, 2 MONOLITHIC */
} covof_momentum_cfg;

/* Unknowns of the momentum system, ns_init (src/module_navier_stokes.f90:
 * 136-226): u(1:u_nx, 1:ny) then v(1:nx, 1:v_ny), i fastest, with u_nx = nx+1
 * (nx if periodic in x) and v_ny = ny+1 (ny if periodic in y), then for
 * solver_type MONOLITHIC p(1:nx, 1:ny); *n = total. */
int covof_momentum_size(covof_ctx *ctx, int solver_type, int *n, int *u_nx, int *v_ny);
/* The fields_old of the momentum block (device, halo layout): the current
 * velocity u, v (linearisation state), u_old, v_old (the time level n),
 * u_nm1, v_nm1 (level n-1; may be NULL unless has_nm1 with GEAR or
 * ADAMS_BASHFORTH), rho_u, rho_v, mu, mu_node, fx_sigma, fy_sigma, and dt.
 * The pointers are kept (not copied) until the next call.  Supported: the
 * standard projection with UNIFORM boundary profiles, no obstacles, no energy
 * equation, no SSP-RK3 momentum, periodic sides in pairs, and no NEUMANN top
 * (the reference's v-row at j = ny+1 reads past its array); anything else
 * returns COVOF_BAD_ARGUMENT. */
int covof_momentum_set_fields_d(covof_ctx *ctx, const covof_momentum_cfg *cfg, const double *u,
                                const double *v, const double *u_old, const double *v_old,
                                const double *u_nm1, const double *v_nm1, const double *rho_u,
                                const double *rho_v, const double *mu, const double *mu_node,
                                const double *fx_sigma, const double *fy_sigma, double dt);
/* diag_u, diag_v (and diag_p, the Schur diagonal, when monolithic) of
 * assemble_preconditioner_blocks (src/module_navier_stokes.f90:1268-1595), the
 * Jacobi preconditioner of the momentum solve; n entries. */
int covof_momentum_diag_d(covof_ctx *ctx, double *diag);
/* assemble_rhs (src/module_navier_stokes.f90:785-1101); n entries. */
int covof_momentum_rhs_d(covof_ctx *ctx, double *rhs);
/* y = A x, apply_system_matrix (src/module_navier_stokes.f90:380-780), with
 * the reference's unpack_vector(homogeneous) into the operator's persistent
 * work fields (its SAVEd tf, zero at the context's first product); when
 * monolithic, the momentum rows carry B^T p and the pressure rows are the
 * continuity rows div(u) dvol. */
int covof_momentum_matvec_d(covof_ctx *ctx, const double *x, double *y);
/* solve_linear_system (src/module_linear_solvers.f90:1167-1208) on the
 * momentum (or monolithic) operator: lin_solver BICGSTAB2 (:361-579) or CG
 * (:735-838), precond NONE, JACOBI (:88-135), SSOR (:137-166) or
 * BLOCK_TRIANGULAR (:168-316), its blocks assembled from the bound fields.
 * x: initial guess in, solution out.  *iters as the reference reports it. */
int covof_momentum_solve_d(covof_ctx *ctx, const double *b, double *x, int lin_solver, int precond,
                           double tol, int maxit, int *iters, double *relres);
/* z = M^-1 r, apply_preconditioner (src/module_linear_solvers.f90:48-316) with
 * the blocks of assemble_preconditioner_blocks of the bound fields: NONE,
 * JACOBI, SSOR (two symmetric lexicographic sweeps per block, :1278-1403) or
 * BLOCK_TRIANGULAR (monolithic: 12 SSOR sweeps at omega 1.25 on the Schur
 * block, mean removed, then 4 sweeps on u and v against its gradient; SSOR
 * otherwise); n entries. */
int covof_momentum_precond_d(covof_ctx *ctx, int precond, const double *r, double *z);
/* Host-array forms, for a CPU solve that keeps its own Krylov loop and hands
 * the preconditioner to the GPU through the reference's only extension point,
 * custom_precond_hook (src/module_linear_solvers.f90:22-41), which
 * apply_preconditioner calls for PRECOND_CUSTOM (:306-312) with
 * (ns, mesh, cfg, src, dst): covof_momentum_set_fields copies the
 * fields_old arrays (host, halo layout) into the context's device mirrors and
 * binds them (they stay bound until another host-array entry point reuses the
 * mirrors); covof_momentum_precond is dst = M^-1 src of
 * covof_momentum_precond_d on host arrays of n entries, n the bound system's
 * size (covof_momentum_size), else COVOF_BAD_ARGUMENT.  The Fortran module
 * procedure covof_gpu_precond (fortran/covof_precond_hook.f90) has the hook's
 * interface and calls it. */
int covof_momentum_set_fields(covof_ctx *ctx, const covof_momentum_cfg *cfg, const double *u, const double *v,
                              const double *u_old, const double *v_old, const double *u_nm1, const double *v_nm1,
                              const double *rho_u, const double *rho_v, const double *mu, const double *mu_node,
                              const double *fx_sigma, const double *fy_sigma, double dt);
int covof_momentum_precond(covof_ctx *ctx, int precond, int n, const double *r, double *z);
/* The momentum block of sim_solve_standard_projection (src/module_simulation.f90:
 * 2454-2469): diagonal, right-hand side, pack (the current u, v are the
 * initial guess), solve, unpack_vector(.false.), sync_u / sync_v and
 * apply_boundary_conditions(.false.).  u, v (halo layout) are updated in place
 * with the predicted velocity u*, v*; the other arguments as
 * covof_momentum_set_fields_d.  Returns the solve's status. */
int covof_sp_predictor_d(covof_ctx *ctx, const covof_momentum_cfg *cfg, double *u, double *v,
                         const double *u_old, const double *v_old, const double *u_nm1,
                         const double *v_nm1, const double *rho_u, const double *rho_v,
                         const double *mu, const double *mu_node, const double *fx_sigma,
                         const double *fy_sigma, double dt, int lin_solver, int precond, double tol,
                         int maxit, int *iters, double *relres);
/* sync_u, sync_v and apply_boundary_conditions(homogeneous) of one rank
 * (src/module_mpi_comm.f90:733-775, :290-537, UNIFORM profiles): the periodic
 * halo exchange of the halo fields u, v (device), then the velocity BC with
 * the Dirichlet values of cfg (zero when homogeneous != 0).  The standard
 * projection calls it after the VOF transport and after the velocity
 * correction (src/module_simulation.f90:541, :2657-2660). */
int covof_velocity_bc_d(covof_ctx *ctx, const covof_momentum_cfg *cfg, double *u, double *v,
                        int homogeneous);
/* Host-array variant of covof_sp_predictor_d (the Fortran drop-in). */
int covof_sp_predictor(covof_ctx *ctx, const covof_momentum_cfg *cfg, double *u, double *v,
                       const double *u_old, const double *v_old, const double *u_nm1,
                       const double *v_nm1, const double *rho_u, const double *rho_v,
                       const double *mu, const double *mu_node, const double *fx_sigma,
                       const double *fy_sigma, double dt, int lin_solver, int precond, double tol,
                       int maxit, int *iters, double *relres);
/* The monolithic branch of the time step (src/module_simulation.f90:553-583,
 * one Uzawa pass; SOLVER_MONOLITHIC, cfg->solver_type must say so):
 * assemble_preconditioner_blocks, assemble_rhs (zero continuity rows), pack
 * (u, v, p: the initial guess), solve_linear_system, unpack_vector(.false.)
 * into u, v, p (pressure exchanged, velocity exchanged with its BC).  The
 * fields as for covof_momentum_set_fields_d; returns as covof_sp_predictor_d. */
int covof_monolithic_step_d(covof_ctx *ctx, const covof_momentum_cfg *cfg, double *u, double *v, double *p,
                            const double *u_old, const double *v_old, const double *u_nm1,
                            const double *v_nm1, const double *rho_u, const double *rho_v,
                            const double *mu, const double *mu_node, const double *fx_sigma,
                            const double *fy_sigma, double dt, int lin_solver, int precond, double tol,
                            int maxit, int *iters, double *relres);
/* The same on host arrays (copied in and out through device mirrors). */
int covof_monolithic_step(covof_ctx *ctx, const covof_momentum_cfg *cfg, double *u, double *v, double *p,
                          const double *u_old, const double *v_old, const double *u_nm1,
                          const double *v_nm1, const double *rho_u, const double *rho_v,
                          const double *mu, const double *mu_node, const double *fx_sigma,
                          const double *fy_sigma, double dt, int lin_solver, int precond, double tol,
                          int maxit, int *iters, double *relres);

/* --- output (src/module_vtk_writer.f90) ------------------------------------ */

/* The quantities of one _history.csv row of vtk_write_diagnostics (:399-531)
 * from the device halo fields C, u, v, p, rho and T (T may be NULL: the
 * reference's reset temperature, 300 everywhere): out[8] (host memory) =
 * center_of_mass_y, rise_velocity, bubble_volume, kinetic_energy, max_u,
 * max_div, avg_rho_gas, avg_p_gas.  Synchronises the context's stream. */
int covof_history_d(covof_ctx *ctx, const double *C, const double *u, const double *v, const double *p,
                    const double *rho, const double *T, double *out);
/* The derived cell fields of vtk_write_step (:90-394): divergence and the
 * cell-centred velocity (uc, vc), (nx, ny) device arrays. */
int covof_cell_output_d(covof_ctx *ctx, const double *u, const double *v, double *div, double *uc,
                        double *vc);

/* --- kernel timing (no reference counterpart: instrumentation) ------------- */

/* Kernel tags of covof_ctx_profile_read. */
#define COVOF_KERNEL_PCG_A_FP64 0 /* fused PCG direction + matvec, FP64 */
#define COVOF_KERNEL_PCG_B_FP64 1 /* fused PCG update + Jacobi + dots, FP64 */
#define COVOF_KERNEL_PCG_A_FP32 2 /* the same kernels of the mixed mode, FP32 */
#define COVOF_KERNEL_PCG_B_FP32 3
#define COVOF_KERNEL_VOF_RECON 4  /* PLIC/ELVIRA reconstruction of one pass */
#define COVOF_KERNEL_VOF_UPDATE 5 /* flux + update + halo of one pass */
#define COVOF_KERNEL_MOM_MATVEC 6 /* the SP momentum operator (one product) */
#define COVOF_NUM_KERNEL_TAGS 7
/* on > 0: time launches of the tagged kernels from their own dispatch start
 * and end timestamps (hipExtLaunchKernelGGL event pairs on the context's
 * stream) and accumulate the durations -- every launch of the VOF kernels
 * and of the momentum operator,
 * the Krylov kernels of every on-th iteration (a sampling period that keeps
 * the bookkeeping out of a timed solve; timed solves run their kernels as
 * direct launches instead of graph replays); on == 0: stop.  Turning it on
 * clears the totals. */
int covof_ctx_profile(covof_ctx *ctx, int on);
/* Accumulated duration (ms) and number of launches of one kernel tag. */
int covof_ctx_profile_read(covof_ctx *ctx, int tag, double *total_ms, long *count);

/* --- adjoint of the pressure projection (no reference counterpart) --------- */

/* Given the forward solution x of A(rho) x = b and w = dJ/dx (nx*ny device
 * arrays, J a scalar objective), solve the transpose system A^T lam = w
 * (A is symmetric) and return dJ/db = lam and dJ/d(log rho) per cell
 * (nx*ny), where rho enters through the face densities rho_u, rho_v of the
 * assembly (arithmetic face means of the cell densities).  rho: cell density
 * (halo layout); dt as in the assembly.  Solver arguments as
 * covof_poisson_solve_d. */
int covof_poisson_adjoint_d(covof_ctx *ctx, const double *rho, double dt, const double *x,
                            const double *w, double *dJ_db, double *dJ_dlogrho, int lin_solver,
                            int precond, double tol, int maxit, int precision, int *iters,
                            double *relres);

/* p(1:nx,1:ny) = x - mean(x) with the periodic halo wrap: the end of the SP
 * Poisson block (src/module_simulation.f90:2577-2588) on a solution x
 * (nx*ny) of covof_poisson_solve_d; p in halo layout. */
int covof_sp_pressure_d(covof_ctx *ctx, const double *x, double *p);

/* Adjoint of the whole pressure projection of the standard projection (no
 * reference counterpart).  Forward (src/module_simulation.f90:2477-2654):
 * b = -div(u*) with the operator A assembled from rho_u, rho_v, dt; A x = b;
 * p = x - mean(x); u = u* - (dt/rho_u) dp/dx, v = v* - (dt/rho_v) dp/dy on
 * the non-wall faces.  The operator must be the forward's (assembled), x its
 * solution (nx*ny, before the mean removal), rho the cell density (halo
 * layout) whose arithmetic face means are rho_u, rho_v.  Given a scalar
 * objective's gradients w_u = dJ/du, w_v = dJ/dv (halo layout) and
 * w_p = dJ/dp (halo layout, NULL: zero; its periodic halo entries count as
 * the copies they are, the others are constants), returns dJ/du*, dJ/dv* (halo layout)
 * and dJ/d(log rho) (nx*ny) through the face means, with halo densities the
 * copies of interior cells (periodic wrap, zero gradient).  One transpose CG
 * solve (A is symmetric) with the given preconditioner and precision.  Not
 * with an obstacle set on the context (COVOF_BAD_ARGUMENT). */
int covof_projection_adjoint_d(covof_ctx *ctx, const double *rho, const double *rho_u, const double *rho_v,
                               double dt, const double *x, const double *w_u, const double *w_v,
                               const double *w_p, double *dJ_du, double *dJ_dv, double *dJ_dlogrho,
                               int precond, double tol, int maxit, int precision, int *iters, double *relres);

/* --- multi-step adjoint of the kinetic energy (no reference counterpart) -- */
/* The reverse-mode pieces of a two-phase step reduced to transport, density
 * update, inertial predictor and projection; a host loop over stored
 * per-step states chains them with covof_projection_adjoint_d into the
 * gradient of the final kinetic energy with respect to the initial C. */

/* Reverse mode of one advect_vof step with WEYMOUTH_YUE (covof_vof_advect_d
 * with nsteps = 1 and the given step_count, which fixes the pass order;
 * two_phase: the physical-corner fix of the halo).  C, u, v: the step's
 * inputs (halo layout).  dJ_dCnew = dJ/dC^{n+1} (halo layout, halo entries
 * included); dJ_dC = dJ/dC^n (halo layout; may be dJ_dCnew itself): with
 * fold_input_halo the halo entries are moved onto the interior cells that
 * C^n's halo copies (zero-gradient copy or wrap, corners from the diagonal
 * cell), else they stay the derivatives with respect to the halo values as
 * given.  dJ_du, dJ_dv: dJ/du, dJ/dv are ADDED to them (face layout).  The
 * map is differentiated with its branches frozen (dilation indicator, donor
 * and empty / mixed / full class of every face and cell, ELVIRA's choice
 * and sign factors, the second pass's clamp), the line constant as the
 * exact root of the volume equation.  Where rounding alone decides a branch
 * the derivative follows a fixed rule: a clamp that removes at most 1e-12
 * passes it, and an ELVIRA winner whose error ties within 1e-10 (relative)
 * with a lower-index candidate of the same unit normal (to 1e-9) is
 * differentiated along that candidate.  Meshes periodic in both directions
 * or in neither, UNIFORM top profile. */
int covof_vof_advect_adjoint_d(covof_ctx *ctx, const double *C, const double *u, const double *v, double dt,
                               const char *reconstruction_method, int step_count, int two_phase,
                               int fold_input_halo, const double *dJ_dCnew, double *dJ_dC, double *dJ_du,
                               double *dJ_dv);

/* C's halo from its interior as the transport leaves it with two_phase:
 * zero-gradient copies on physical sides (apply_vof_bc,
 * src/module_mpi_comm.f90:636-683), the wrap on periodic ones, the corners
 * from the diagonal cell (src/module_interface_vof.f90:532-535).  Meshes
 * periodic in both directions or in neither. */
int covof_vof_fill_halo_d(covof_ctx *ctx, double *C);

/* Kinetic energy of the history row (src/module_vtk_writer.f90:399-531),
 * KE = sum 0.5 rho (0.5 (uw^2 + ue^2) + 0.5 (vs^2 + vn^2)) dx dy with rho
 * from the clamped C as update_physical_properties forms it; into *ke (host
 * memory, synchronises; NULL: skipped), and its gradients dKE/dC, dKE/du,
 * dKE/dv (halo layout, overwritten; each may be NULL). */
int covof_kinetic_energy_d(covof_ctx *ctx, const double *C, const double *u, const double *v, double rho1,
                           double rho2, double *ke, double *dKE_dC, double *dKE_du, double *dKE_dv);

/* Reverse mode of the density of update_physical_properties
 * (src/module_interface_vof.f90:552-553): dJ_dC(interior) += dJ/d(log rho)
 * (nx*ny) * (rho1 - rho2) / rho where the clamp passes C through. */
int covof_properties_adjoint_d(covof_ctx *ctx, const double *C, double rho1, double rho2,
                               const double *dJ_dlogrho, double *dJ_dC);

/* u* of the momentum block with convection and viscosity off: the
 * diagonal alpha rho / dt (alpha = 1.5 with BDF2) and the right-hand side
 * rho (2 u^n - 0.5 u^{n-1}) / dt + rho g (src/module_navier_stokes.f90:
 * 841-847, :895-904) give u* = (4 u^n - u^{n-1} + 2 dt g) / 3, or
 * u^n + dt g when first (explicit Euler inertia); zero on wall
 * faces and outside the face ranges.  Every side a wall (DIRICHLET,
 * FREESLIP) or periodic.  u_nm1, v_nm1 may be NULL when first. */
int covof_inertial_predictor_d(covof_ctx *ctx, int first, double dt, double gx, double gy, const double *u,
                               const double *v, const double *u_nm1, const double *v_nm1, double *u_star,
                               double *v_star);
/* Its transpose: dJ_du += a w_u, dJ_du_nm1 += b w_u (and v) on the
 * predictor's faces, a = 4/3, b = -1/3 (first: a = 1, no n-1 term). */
int covof_inertial_predictor_adjoint_d(covof_ctx *ctx, int first, const double *w_u, const double *w_v,
                                       double *dJ_du, double *dJ_dv, double *dJ_du_nm1, double *dJ_dv_nm1);

#ifdef __cplusplus
}
#endif
#endif /* COVOF_HIP_H */
