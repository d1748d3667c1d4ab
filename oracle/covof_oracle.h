/*
 * CPU oracle of the COVOF hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline).
 * The product path never calls it.
 *
 * Layout conventions (the reference's, src/module_fields.f90:129-169):
 *   halo fields  (0:nx+1, 0:ny+1) column-major, i fastest:  F[j*(nx+2)+i]
 *   Poisson work arrays (1:nx, 1:ny) column-major:         A[(j-1)*nx+(i-1)]
 * Boundary types are the reference's integers (src/module_constants.f90:25-31):
 *   1 DIRICHLET, 2 NEUMANN, 3 PERIODIC, 4 FREESLIP.
 */
#ifndef COVOF_ORACLE_H
#define COVOF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny;          /* cells (single rank: local == global) */
  double lx, ly, dx, dy;
  int bc[4];           /* left, right, bottom, top */
} orc_mesh;

void orc_mesh_init(orc_mesh *m, int nx, int ny, double lx, double ly, const int bc[4]);

double orc_cut_box_area(double xmin, double xmax, double ymin, double ymax,
                        double nx, double ny, double alpha);
double orc_compute_alpha(double xmin, double xmax, double ymin, double ymax,
                         double nx, double ny, double c_target);

/* method: 0 = PLIC (Youngs), 1 = ELVIRA */
void orc_reconstruct(const orc_mesh *m, int method, const double *C,
                     double *snx, double *sny, double *sal);
void orc_reconstruct_interface(const orc_mesh *m, int method, const double *C,
                               double *snx, double *sny, double *sal);
void orc_exchange_halos_scalar(const orc_mesh *m, double *f);
void orc_apply_vof_bc(const orc_mesh *m, double *C);

/* One advect_vof(WEYMOUTH_YUE) call: both passes, step counter, halo, BC.
 * If two_phase, also update_physical_properties (rho, rho_u, rho_v, mu, mu_node;
 * any of those may be NULL except that all are needed when two_phase != 0). */
void orc_wy_step(const orc_mesh *m, int method, double *C, const double *u, const double *v,
                 double dt, int *step_count);
void orc_update_properties(const orc_mesh *m, double *C, double rho1, double rho2,
                           double mu1, double mu2, double *rho, double *rho_u,
                           double *rho_v, double *mu, double *mu_node);

/* The steps advect_vof runs after the transport (single rank, no markers):
 * SVOF Helmholtz smoothing C -> Cs (src/module_interface_vof.f90:387-476),
 * diffusion smoothing C -> Ct (:1416-1461; steps < 1 means 3), and the
 * curvature + CSF forces (:627-789) with the normals taken from G (C, C_s or
 * C_tilde) and the force gradient from C; enable == 0 or sigma <= 0 zeroes
 * curv, fx, fy.  All arrays are halo fields. */
void orc_svof_smooth(const orc_mesh *m, const double *C, double width, double *Cs);
void orc_smooth_diffusion(const orc_mesh *m, const double *C, int steps, double *Ct);
/* method 0: the reference's curvature; 1: height-function curvature where the
 * heights bracket the interface (no reference counterpart) */
void orc_curvature_csf2(const orc_mesh *m, const double *G, const double *C, int enable, double sigma,
                        int method, double *curv, double *fx, double *fy);
int orc_hf_kappa(const orc_mesh *m, const double *C, int i, int j, double nxv, double nyv, double *kap);
void orc_curvature_csf(const orc_mesh *m, const double *G, const double *C, int enable, double sigma,
                       double *curv, double *fx, double *fy);

/* Poisson (SP projection block). Coefficient arrays are (nx,ny). */
void orc_poisson_assemble(const orc_mesh *m, const double *u, const double *v,
                          const double *rho_u, const double *rho_v, double dt,
                          double *aW, double *aE, double *aS, double *aN, double *aP,
                          double *b);
void orc_poisson_matvec(const orc_mesh *m, const double *aP, const double *aW,
                        const double *aE, const double *aS, const double *aN,
                        const double *x, double *y);
void orc_poisson_sgs(const orc_mesh *m, const double *aP, const double *aW,
                     const double *aE, const double *aS, const double *aN,
                     const double *r, double *z);
/* precond: 0 none, 1 Jacobi, 2 SGS ("SSOR" omega=1) */
void orc_poisson_precond(const orc_mesh *m, int precond, const double *aP, const double *aW,
                         const double *aE, const double *aS, const double *aN,
                         const double *r, double *z);
/* SP PCG driver (src/module_simulation.f90:2535-2585) on x (in/out, (nx,ny)).
 * Returns iterations; *relres = final ||r||/||b||.  No mean removal here. */
int orc_sp_pcg(const orc_mesh *m, int precond, const double *aP, const double *aW,
               const double *aE, const double *aS, const double *aN,
               const double *b, double *x, double tol, int maxit, double *relres);
/* BiCGSTAB(2) of src/module_linear_solvers.f90:361-579 on the Poisson operator. */
int orc_bicgstab2(const orc_mesh *m, int precond, const double *aP, const double *aW,
                  const double *aE, const double *aS, const double *aN,
                  const double *b, double *x, double tol, int maxit, double *relres);
/* Mean removal (src/module_simulation.f90:2578-2585) into p's interior, then
 * sync_cell_scalar of the single-rank case (periodic wrap). */
void orc_sp_finish_pressure(const orc_mesh *m, const double *x, double *p);
void orc_sp_correct(const orc_mesh *m, const double *p, const double *rho_u,
                    const double *rho_v, double dt, double *u, double *v);

/* SP velocity predictor (src/module_navier_stokes.f90, src/module_linear_solvers.f90,
 * src/module_simulation.f90:2454-2469).  Supported subset: standard projection,
 * no obstacles, UNIFORM boundary profiles, no energy equation, no SSP-RK3.
 * Unknown vector: u(1:u_nx,1:ny) then v(1:nx,1:v_ny), i fastest, with
 * u_nx = nx + 1 (nx if periodic in x) and v_ny = ny + 1 (ny if periodic in y). */
typedef struct {
  int time_scheme;        /* 1 EULER, 2 GEAR (BDF2) */
  int has_nm1;            /* fields%has_nm1 */
  int enable_convection;
  int conv_scheme;        /* 1 UPWIND, 2 CENTERED, 3 TVD_SUPERBEE */
  int conv_lin;           /* 1 PICARD, 2 ADAMS_BASHFORTH, 3 NEWTON */
  double gx, gy;
  double bc_u[4], bc_v[4];  /* Dirichlet values: left, right, bottom, top */
  int solver_type;        /* 3 STANDARD_PROJECTION (u, v unknowns), 2 MONOLITHIC (u, v, p) */
} orc_mom_cfg;
/* Outputs of assemble_preconditioner_blocks: diag (n), u_a* (nu), v_a* (nv),
 * schur_a* (nx*ny); any may be NULL. */
typedef struct {
  double *diag, *uW, *uE, *uS, *uN, *uP, *vW, *vE, *vS, *vN, *vP, *sW, *sE, *sS, *sN, *sP;
} orc_mom_blocks_t;
typedef struct {
  const double *u, *v, *u_old, *v_old, *u_nm1, *v_nm1;
  const double *rho_u, *rho_v, *mu, *mu_node, *fx, *fy;  /* fx/fy: fx_sigma, fy_sigma */
} orc_mom_fields;
void orc_velocity_bc(const orc_mesh *m, const orc_mom_cfg *c, double *u, double *v, int homogeneous);
void orc_mom_sizes(const orc_mesh *m, int *u_nx, int *v_ny);
void orc_mom_pack(const orc_mesh *m, const double *u, const double *v, double *vec);
void orc_mom_unpack(const orc_mesh *m, const orc_mom_cfg *c, const double *vec, double *u, double *v,
                    int homogeneous);
/* tu, tv: persistent halo work fields (the reference's SAVEd `tf`), zero at first use. */
void orc_mom_matvec(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                    double *tu, double *tv, const double *x, double *y);
void orc_mom_diag(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                  double *diag);
void orc_mom_rhs(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                 double *rhs);
/* lin_solver 1 BICGSTAB2, 3 CG; precond 0 NONE, 1 JACOBI.  Returns iters. */
int orc_mom_solve(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                  const double *diag, double *tu, double *tv, int lin_solver, int precond, double tol,
                  int maxit, const double *b, double *x, double *relres);
int orc_sp_predictor(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                     double *tu, double *tv, int lin_solver, int precond, double tol, int maxit,
                     double *u, double *v, double *relres);
/* The monolithic system (solver_type 2): unknowns u, v, then p(1:nx,1:ny). */
void orc_mono_pack(const orc_mesh *m, const double *u, const double *v, const double *p, double *vec);
void orc_mono_unpack(const orc_mesh *m, const orc_mom_cfg *c, const double *vec, double *u, double *v,
                     double *p, int homogeneous);
void orc_mono_matvec(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                     double *tu, double *tv, double *tp, const double *x, double *y);
void orc_mom_blocks(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                    const orc_mom_blocks_t *B);
void orc_ssor_block(const orc_mesh *m, int ni, int nj, const double *aW, const double *aE, const double *aS,
                    const double *aN, const double *aP, const double *rhs, double *x, int n_sweeps,
                    double omega, int sync);
/* precond 0 NONE, 1 JACOBI, 2 SSOR, 4 BLOCK_TRIANGULAR */
void orc_mom_precond(const orc_mesh *m, const orc_mom_cfg *c, int precond, const orc_mom_blocks_t *B,
                     const double *src, double *dst);
int orc_mom_solve2(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt, double *tu,
                   double *tv, double *tp, int lin_solver, int precond, double tol, int maxit, const double *b,
                   double *x, double *relres);
int orc_mono_step(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt, double *tu,
                  double *tv, double *tp, int lin_solver, int precond, double tol, int maxit, double *u, double *v,
                  double *p, double *relres);

/* vtk_write_diagnostics (src/module_vtk_writer.f90:399-531), serial sums:
 * out[8] = center_of_mass_y, rise_velocity, bubble_volume, kinetic_energy,
 * max_u, max_div, avg_rho_gas, avg_p_gas; T may be NULL (300 everywhere). */
void orc_history(const orc_mesh *m, const double *C, const double *u, const double *v, const double *p,
                 const double *rho, const double *T, double *out);

/* Serial column-major sum of x*y (the reference's sum(x*y)). */
double orc_dot(const double *x, const double *y, long n);

/* Timed CPU baseline: `iters` Jacobi-PCG iterations on the face-form
 * operator with OpenMP over `threads` threads (FP64). Returns seconds. */
void orc_set_threads(int threads);
void orc_set_obstacle(int enabled, int type, double x0, double y0, double x1, double y1, double xc, double yc,
                      double r, double perm);
void orc_set_top_jet(int on, double center, double width);
void orc_update_properties_cs(const orc_mesh *m, double *C, const double *Cs, double rho1, double rho2,
                              double mu1, double mu2, double *rho, double *rho_u, double *rho_v, double *mu,
                              double *mu_node);
double orc_bench_jacobi_pcg_solve(const orc_mesh *m, const double *aP, const double *aW,
                                  const double *aE, const double *aS, const double *aN,
                                  const double *b, double tol, int maxit, int threads, int *iters,
                                  double *relres);
double orc_bench_jacobi_pcg(const orc_mesh *m, const double *aP, const double *aW,
                            const double *aE, const double *aS, const double *aN,
                            const double *b, int iters, int threads);

#ifdef __cplusplus
}
#endif
#endif
