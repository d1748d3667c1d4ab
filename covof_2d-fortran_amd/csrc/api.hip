// extern "C" entry points of libcovof_hip (declared in include/covof_hip.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "covof_internal.h"

namespace covof {
int vof_advect(Ctx &c, double *C, const double *u, const double *v, double dt, int method,
               int *step_count, int nsteps, int two_phase, double rho1, double rho2, double mu1,
               double mu2, double *rho, double *rho_u, double *rho_v, double *mu, double *mu_node);
int vof_properties(Ctx &c, const double *C, double rho1, double rho2, double mu1, double mu2,
                   double *rho, double *rho_u, double *rho_v, double *mu, double *mu_node, const double *Cs);
int vof_reconstruct_interface(Ctx &c, int method, const double *C, double *onx, double *ony,
                              double *oal);
int vof_geom_batch(Ctx &c, int n, const double *xmin, const double *xmax, const double *ymin,
                   const double *ymax, const double *gnx, const double *gny, const double *alpha,
                   const double *cc, double *area, double *aout);
int vof_corner_fix(Ctx &c, double *C);
int noalloc_scratch(int **out);
int noalloc_max_tiles();
int noalloc_tiles(const Mesh &m);
int vof_reconstruct_noalloc(const Mesh &mesh, int method, const double *C, double *onx, double *ony, double *oal,
                            hipStream_t s);
int poisson_assemble(Ctx &c, const double *u, const double *v, const double *rho_u,
                     const double *rho_v, double dt, double *b);
int poisson_matvec(Ctx &c, const double *x, double *y);
int poisson_precond(Ctx &c, int precond, const double *r, double *z);
int poisson_solve(Ctx &c, const double *b, double *x, int lin_solver, int precond, double tol,
                  int maxit, int precision, int *iters, double *relres);
int sp_poisson(Ctx &c, const double *u, const double *v, const double *rho_u, const double *rho_v,
               double *p, double dt, int lin_solver, int precond, double tol, int maxit,
               int precision, int *iters, double *relres);
int sp_correct(Ctx &c, const double *p, const double *rho_u, const double *rho_v, double dt,
               double *u, double *v);
int poisson_adjoint(Ctx &c, const double *rho, double dt, const double *x, const double *w,
                    double *dJ_db, double *dJ_dlogrho, int lin_solver, int precond, double tol,
                    int maxit, int precision, int *iters, double *relres);
int projection_adjoint(Ctx &c, const double *rho, const double *rho_u, const double *rho_v, double dt,
                       const double *x, const double *w_u, const double *w_v, const double *w_p, double *dJ_du,
                       double *dJ_dv, double *dJ_dlogrho, int precond, double tol, int maxit, int precision,
                       int *iters, double *relres);
int vof_advect_adjoint(Ctx &c, int method, const double *C, const double *u, const double *v, double dt,
                       int step_count, int two_phase, int fold_in, const double *cbar_new, double *cbar,
                       double *ubar, double *vbar);
int vof_fill_halo(Ctx &c, double *C);
int kinetic_energy(Ctx &c, const double *C, const double *u, const double *v, double rho1, double rho2, double *ke,
                   double *dC, double *du, double *dv);
int properties_adjoint(Ctx &c, const double *C, double rho1, double rho2, const double *g, double *dC);
int inertial_predictor(Ctx &c, int first, double dt, double gx, double gy, const double *u, const double *v,
                       const double *un1, const double *vn1, double *us, double *vs);
int inertial_predictor_adjoint(Ctx &c, int first, const double *wu, const double *wv, double *du, double *dv,
                               double *dun1, double *dvn1);
}  // namespace covof

using covof::Ctx;

struct covof_ctx {
  Ctx c;
  // device mirrors for the host-array entry points (made on first use)
  double *mirror[13] = {nullptr};
  // r and z of covof_momentum_precond (made on first use, monolithic size)
  double *hvec[2] = {nullptr};
};

static thread_local std::string g_err;

static int fail(int code, const char *msg) {
  g_err = msg;
  return code;
}

static int hip_status(int st) {
  if (st == COVOF_HIP_ERROR) {
    hipError_t e = hipGetLastError();
    g_err = std::string("HIP error: ") + hipGetErrorString(e);
  }
  return st;
}

extern "C" const char *covof_last_error(void) { return g_err.c_str(); }

// The mesh of a context-free entry point, with covof_ctx_create's checks.
static int mesh_of(int nx, int ny, double lx, double ly, const int *bc, covof::Mesh *m) {
  if (!bc || nx < 2 || ny < 2 || !(lx > 0.0) || !(ly > 0.0)) return COVOF_BAD_ARGUMENT;
  if ((long)(nx + 2) * (ny + 2) > 0x7fffffffL) return COVOF_BAD_ARGUMENT;
  for (int k = 0; k < 4; ++k)
    if (bc[k] < 1 || bc[k] > 4) return COVOF_BAD_ARGUMENT;
  m->nx = nx;
  m->ny = ny;
  m->lx = lx;
  m->ly = ly;
  m->dx = lx / (double)nx;
  m->dy = ly / (double)ny;
  for (int k = 0; k < 4; ++k) m->bc[k] = bc[k];
  m->perx = bc[0] == covof::BC_PERIODIC || bc[1] == covof::BC_PERIODIC;
  m->pery = bc[2] == covof::BC_PERIODIC || bc[3] == covof::BC_PERIODIC;
  return COVOF_OK;
}

// "WEYMOUTH_YUE" is the only geometric advection scheme of this library; the
// reconstruction name follows reconstruct_interface's select case
// (src/module_interface_reconstruction.f90:591-600): "ELVIRA" -> ELVIRA,
// "SLIC"/"LVIRA" are not provided, anything else -> PLIC.
static int recon_code(const char *name, int *code) {
  if (!name) return COVOF_BAD_ARGUMENT;
  std::string s(name);
  while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
  if (s == "ELVIRA") *code = 1;
  else if (s == "SLIC" || s == "LVIRA") return COVOF_BAD_ARGUMENT;
  else *code = 0;
  return COVOF_OK;
}

static bool is_wy(const char *name) {
  if (!name) return false;
  std::string s(name);
  while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
  return s == "WEYMOUTH_YUE";
}

extern "C" int covof_ctx_create(int nx, int ny, double lx, double ly, const int *bc, covof_ctx **out) {
  if (!out || !bc || nx < 2 || ny < 2 || !(lx > 0.0) || !(ly > 0.0))
    return fail(COVOF_BAD_ARGUMENT, "covof_ctx_create: need nx, ny >= 2, lx, ly > 0, bc[4]");
  if ((long)(nx + 2) * (ny + 2) > 0x7fffffffL)
    return fail(COVOF_BAD_ARGUMENT, "covof_ctx_create: grid too large for 32-bit cell indices");
  for (int k = 0; k < 4; ++k)
    if (bc[k] < 1 || bc[k] > 4) return fail(COVOF_BAD_ARGUMENT, "covof_ctx_create: bc type must be 1..4");
  covof_ctx *h = new (std::nothrow) covof_ctx();
  if (!h) return fail(COVOF_BAD_ARGUMENT, "out of host memory");
  Ctx &c = h->c;
  c.m.nx = nx;
  c.m.ny = ny;
  c.m.lx = lx;
  c.m.ly = ly;
  c.m.dx = lx / (double)nx;
  c.m.dy = ly / (double)ny;
  for (int k = 0; k < 4; ++k) c.m.bc[k] = bc[k];
  c.m.perx = bc[0] == covof::BC_PERIODIC || bc[1] == covof::BC_PERIODIC;
  c.m.pery = bc[2] == covof::BC_PERIODIC || bc[3] == covof::BC_PERIODIC;
  if (hipStreamCreate(&c.stream) != hipSuccess) {
    delete h;
    return hip_status(COVOF_HIP_ERROR);
  }
  c.own_stream = 1;
  // the device address of the allocation-free entry point's scratch array on
  // this GPU (looked up once per device)
  int *scratch_dev = nullptr;
  int st = covof::noalloc_scratch(&scratch_dev);
  if (!st) st = covof::vof_alloc(c);
  if (!st) st = covof::poisson_alloc(c);
  if (st) {
    covof_ctx_destroy(h);
    return hip_status(st);
  }
  *out = h;
  return COVOF_OK;
}

extern "C" int covof_ctx_destroy(covof_ctx *h) {
  if (!h) return COVOF_OK;
  Ctx &c = h->c;
  if (c.stream) hipStreamSynchronize(c.stream);
  covof::vof_free(c);
  covof::surface_free(c);
  covof::poisson_free(c);
  covof::mom_free(c);
  if (c.prof.created)
    for (int k = 0; k < 2 * 1024; ++k) hipEventDestroy(c.prof.ev[k]);
  for (double *p : h->mirror) hipFree(p);
  for (double *p : h->hvec) hipFree(p);
  if (c.own_stream && c.stream) hipStreamDestroy(c.stream);
  delete h;
  return COVOF_OK;
}

extern "C" int covof_ctx_set_stream(covof_ctx *h, void *stream) {
  if (!h) return fail(COVOF_BAD_ARGUMENT, "null context");
  Ctx &c = h->c;
  if (stream) {
    if (c.own_stream) {
      hipStreamSynchronize(c.stream);
      hipStreamDestroy(c.stream);
      c.own_stream = 0;
    }
    c.stream = (hipStream_t)stream;
  } else if (!c.own_stream) {
    if (hipStreamCreate(&c.stream) != hipSuccess) return hip_status(COVOF_HIP_ERROR);
    c.own_stream = 1;
  }
  return COVOF_OK;
}

extern "C" int covof_ctx_set_reduction_order(covof_ctx *h, int order) {
  if (!h) return fail(COVOF_BAD_ARGUMENT, "null context");
  if (order != COVOF_REDUCTION_TREE && order != COVOF_REDUCTION_REFERENCE)
    return fail(COVOF_BAD_ARGUMENT, "covof_ctx_set_reduction_order: order must be TREE (0) or REFERENCE (1)");
  h->c.ref_order = order == COVOF_REDUCTION_REFERENCE;
  return COVOF_OK;
}

extern "C" int covof_ctx_set_solver_option(covof_ctx *h, int option, double value) {
  if (!h) return fail(COVOF_BAD_ARGUMENT, "null context");
  covof::SolverOpts &o = h->c.opts;
  const bool whole = value == (double)(int)value;
  switch (option) {
    case COVOF_OPT_MIXED_STALL_WINDOW:
      if (!whole || value < 0.0) return fail(COVOF_BAD_ARGUMENT, "MIXED_STALL_WINDOW must be a whole number >= 0");
      o.mixed_stall = (int)value;
      break;
    case COVOF_OPT_MIXED_CYCLE:
      if (!whole || value < 0.0 || value == 1.0 || ((int)value & 1))
        return fail(COVOF_BAD_ARGUMENT, "MIXED_CYCLE must be 0 or an even whole number >= 2");
      o.mixed_cycle = (int)value;
      break;
    case COVOF_OPT_MIXED_LOOP:
      if (value != 0.0 && value != 1.0) return fail(COVOF_BAD_ARGUMENT, "MIXED_LOOP must be 0 or 1");
      o.mixed_loop = (int)value;
      break;
    case COVOF_OPT_MG_COARSE:
      if (value != 0.0 && value != 1.0) return fail(COVOF_BAD_ARGUMENT, "MG_COARSE must be 0 or 1");
      o.mg_coarse = (int)value;
      break;
    case COVOF_OPT_MG_SWEEPS:
      if (!whole || value < 1.0 || value > 4.0) return fail(COVOF_BAD_ARGUMENT, "MG_SWEEPS must be 1, 2, 3 or 4");
      o.mg_sweeps = (int)value;
      break;
    case COVOF_OPT_MG_OMEGA:
      if (!(value > 0.0 && value <= 2.0)) return fail(COVOF_BAD_ARGUMENT, "MG_OMEGA must be in (0, 2]");
      o.mg_omega = value;
      break;
    case COVOF_OPT_MG_FUSED:
      if (value != 0.0 && value != 1.0) return fail(COVOF_BAD_ARGUMENT, "MG_FUSED must be 0 or 1");
      o.mg_fused = (int)value;
      break;
    case COVOF_OPT_MIXED_MG_FORM:
      if (value != 0.0 && value != 1.0) return fail(COVOF_BAD_ARGUMENT, "MIXED_MG_FORM must be 0 or 1");
      o.mixed_mg = (int)value;
      break;
    case COVOF_OPT_MG_PROJECT:
      if (value != 0.0 && value != 1.0) return fail(COVOF_BAD_ARGUMENT, "MG_PROJECT must be 0 or 1");
      o.mg_project = (int)value;
      break;
    default:
      return fail(COVOF_BAD_ARGUMENT, "covof_ctx_set_solver_option: unknown option");
  }
  if (hipStreamSynchronize(h->c.stream) != hipSuccess) return hip_status(COVOF_HIP_ERROR);
  covof::solver_invalidate(h->c);
  return COVOF_OK;
}

static double *mirror(covof_ctx *h, int k);

extern "C" int covof_ctx_reserve(covof_ctx *h, int parts) {
  if (!h) return fail(COVOF_BAD_ARGUMENT, "null context");
  if (parts & ~COVOF_RESERVE_ALL) return fail(COVOF_BAD_ARGUMENT, "covof_ctx_reserve: unknown part bits");
  Ctx &c = h->c;
  int st = COVOF_OK;
  if (!st && (parts & COVOF_RESERVE_MOMENTUM)) {
    st = covof::mom_reserve(c);
  }
  if (!st && (parts & COVOF_RESERVE_MULTIGRID)) st = covof::mg_reserve(c);
  if (!st && (parts & COVOF_RESERVE_SGS)) st = covof::sgs_reserve(c);
  if (!st && (parts & COVOF_RESERVE_MIXED_LOOP)) st = covof::cg2_reserve(c);
  if (!st && (parts & COVOF_RESERVE_SURFACE)) st = covof::surface_reserve(c);
  if (!st && (parts & COVOF_RESERVE_ADJOINT)) st = covof::vof_adjoint_reserve(c);
  if (!st && (parts & COVOF_RESERVE_HOST_MIRRORS))
    for (int k = 0; k < 13 && !st; ++k)
      if (!mirror(h, k)) st = COVOF_HIP_ERROR;
  if (!st && hipStreamSynchronize(c.stream) != hipSuccess) st = COVOF_HIP_ERROR;
  return hip_status(st);
}

extern "C" int covof_ctx_mixed_solve_info(covof_ctx *h, int *fp32_iters, int *fp64_iters) {
  if (!h) return fail(COVOF_BAD_ARGUMENT, "null context");
  if (fp32_iters) *fp32_iters = h->c.ps.last_it32;
  if (fp64_iters) *fp64_iters = h->c.ps.last_it64;
  return COVOF_OK;
}

extern "C" int covof_ctx_synchronize(covof_ctx *h) {
  if (!h) return fail(COVOF_BAD_ARGUMENT, "null context");
  if (hipStreamSynchronize(h->c.stream) != hipSuccess) return hip_status(COVOF_HIP_ERROR);
  return COVOF_OK;
}

extern "C" int covof_geom_batch_d(covof_ctx *h, int n, const double *xmin, const double *xmax,
                                  const double *ymin, const double *ymax, const double *gnx,
                                  const double *gny, const double *alpha, const double *c,
                                  double *area, double *alpha_out) {
  if (!h || n < 0) return fail(COVOF_BAD_ARGUMENT, "covof_geom_batch_d: bad argument");
  return hip_status(covof::vof_geom_batch(h->c, n, xmin, xmax, ymin, ymax, gnx, gny, alpha, c, area, alpha_out));
}

extern "C" int covof_reconstruct_interface_d(covof_ctx *h, const char *method, const double *C,
                                             double *seg_nx, double *seg_ny, double *seg_alpha) {
  int code;
  if (!h || !C || !seg_nx || !seg_ny || !seg_alpha) return fail(COVOF_BAD_ARGUMENT, "null argument");
  if (recon_code(method, &code)) return fail(COVOF_BAD_ARGUMENT, "reconstruction method must be PLIC or ELVIRA");
  return hip_status(covof::vof_reconstruct_interface(h->c, code, C, seg_nx, seg_ny, seg_alpha));
}

extern "C" int covof_reconstruct_interface_noalloc_d(int nx, int ny, double lx, double ly, const int *bc,
                                                     const char *method, const double *C, double *seg_nx,
                                                     double *seg_ny, double *seg_alpha, void *stream) {
  covof::Mesh m;
  int code;
  if (mesh_of(nx, ny, lx, ly, bc, &m))
    return fail(COVOF_BAD_ARGUMENT, "covof_reconstruct_interface_noalloc_d: need nx, ny >= 2, lx, ly > 0, bc[4]");
  if (!C || !seg_nx || !seg_ny || !seg_alpha) return fail(COVOF_BAD_ARGUMENT, "null argument");
  if (seg_nx == seg_ny || seg_nx == seg_alpha || seg_ny == seg_alpha || C == seg_nx || C == seg_ny ||
      C == seg_alpha)
    return fail(COVOF_BAD_ARGUMENT, "covof_reconstruct_interface_noalloc_d: outputs must be distinct from C and "
                                    "from each other");
  if (recon_code(method, &code)) return fail(COVOF_BAD_ARGUMENT, "reconstruction method must be PLIC or ELVIRA");
  if (covof::noalloc_tiles(m) > covof::noalloc_max_tiles())
    return fail(COVOF_BAD_ARGUMENT, (std::string("covof_reconstruct_interface_noalloc_d: the mesh needs ") +
                                     std::to_string(covof::noalloc_tiles(m)) + " tiles of 64 x 16 cells, more than the " +
                                     std::to_string(covof::noalloc_max_tiles()) + " its scratch holds")
                                        .c_str());
  return hip_status(covof::vof_reconstruct_noalloc(m, code, C, seg_nx, seg_ny, seg_alpha, (hipStream_t)stream));
}

extern "C" int covof_vof_advect_d(covof_ctx *h, double *C, const double *u, const double *v, double dt,
                                  const char *advection_method, const char *reconstruction_method,
                                  int *step_count, int nsteps, int two_phase, double rho1, double rho2,
                                  double mu1, double mu2, double *rho, double *rho_u, double *rho_v,
                                  double *mu, double *mu_node) {
  int code;
  if (!h || !C || !u || !v || !step_count || nsteps < 0) return fail(COVOF_BAD_ARGUMENT, "null argument");
  if (!is_wy(advection_method)) return fail(COVOF_BAD_ARGUMENT, "advection method must be WEYMOUTH_YUE");
  if (recon_code(reconstruction_method, &code)) return fail(COVOF_BAD_ARGUMENT, "reconstruction method must be PLIC or ELVIRA");
  if (two_phase && (!rho || !rho_u || !rho_v || !mu || !mu_node))
    return fail(COVOF_BAD_ARGUMENT, "two_phase needs rho, rho_u, rho_v, mu, mu_node");
  if (!(dt >= 0.0)) return fail(COVOF_BAD_ARGUMENT, "dt must be >= 0");
  return hip_status(covof::vof_advect(h->c, C, u, v, dt, code, step_count, nsteps, two_phase, rho1, rho2,
                                      mu1, mu2, rho, rho_u, rho_v, mu, mu_node));
}

// Host mirrors: slot k holds a halo-size device array.
static double *mirror(covof_ctx *h, int k) {
  if (!h->mirror[k]) {
    const long nh = (long)(h->c.m.nx + 2) * (h->c.m.ny + 2);
    if (hipMalloc(&h->mirror[k], nh * sizeof(double)) != hipSuccess) return nullptr;
  }
  return h->mirror[k];
}

static int h2d(covof_ctx *h, double *d, const double *s) {
  const long nh = (long)(h->c.m.nx + 2) * (h->c.m.ny + 2);
  return hipMemcpyAsync(d, s, nh * sizeof(double), hipMemcpyHostToDevice, h->c.stream) == hipSuccess ? 0 : COVOF_HIP_ERROR;
}
static int d2h(covof_ctx *h, double *d, const double *s) {
  const long nh = (long)(h->c.m.nx + 2) * (h->c.m.ny + 2);
  return hipMemcpyAsync(d, s, nh * sizeof(double), hipMemcpyDeviceToHost, h->c.stream) == hipSuccess ? 0 : COVOF_HIP_ERROR;
}

extern "C" int covof_vof_advect(covof_ctx *h, double *C, const double *u, const double *v, double dt,
                                const char *advection_method, const char *reconstruction_method,
                                int *step_count, int nsteps, int two_phase, double rho1, double rho2,
                                double mu1, double mu2, double *rho, double *rho_u, double *rho_v,
                                double *mu, double *mu_node) {
  if (!h || !C || !u || !v) return fail(COVOF_BAD_ARGUMENT, "null argument");
  double *d[8];
  for (int k = 0; k < 8; ++k)
    if (!(d[k] = mirror(h, k))) return hip_status(COVOF_HIP_ERROR);
  int st = h2d(h, d[0], C);
  if (!st) st = h2d(h, d[1], u);
  if (!st) st = h2d(h, d[2], v);
  if (!st && two_phase) {
    if (!rho || !rho_u || !rho_v || !mu || !mu_node) return fail(COVOF_BAD_ARGUMENT, "two_phase needs property arrays");
    const double *hp[5] = {rho, rho_u, rho_v, mu, mu_node};
    for (int k = 0; k < 5 && !st; ++k) st = h2d(h, d[3 + k], hp[k]);
  }
  if (st) return hip_status(st);
  st = covof_vof_advect_d(h, d[0], d[1], d[2], dt, advection_method, reconstruction_method, step_count,
                          nsteps, two_phase, rho1, rho2, mu1, mu2, d[3], d[4], d[5], d[6], d[7]);
  if (st) return st;
  st = d2h(h, C, d[0]);
  if (!st && two_phase) {
    double *hp[5] = {rho, rho_u, rho_v, mu, mu_node};
    for (int k = 0; k < 5 && !st; ++k) st = d2h(h, hp[k], d[3 + k]);
  }
  if (!st && hipStreamSynchronize(h->c.stream) != hipSuccess) st = COVOF_HIP_ERROR;
  return hip_status(st);
}

extern "C" int covof_update_properties_d(covof_ctx *h, double *C, double rho1, double rho2, double mu1,
                                         double mu2, double *rho, double *rho_u, double *rho_v, double *mu,
                                         double *mu_node) {
  if (!h || !C || !rho || !rho_u || !rho_v || !mu || !mu_node) return fail(COVOF_BAD_ARGUMENT, "null argument");
  int st = covof::vof_corner_fix(h->c, C);
  if (!st) st = covof::vof_properties(h->c, C, rho1, rho2, mu1, mu2, rho, rho_u, rho_v, mu, mu_node, nullptr);
  return hip_status(st);
}

extern "C" int covof_update_properties_cs_d(covof_ctx *h, double *C, const double *C_s, double rho1, double rho2,
                                            double mu1, double mu2, double *rho, double *rho_u, double *rho_v,
                                            double *mu, double *mu_node) {
  if (!h || !C || !rho || !rho_u || !rho_v || !mu || !mu_node) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const double *outs[5] = {rho, rho_u, rho_v, mu, mu_node};
  for (const double *o : outs)
    if (o == C || (C_s && o == C_s))
      return fail(COVOF_BAD_ARGUMENT, "covof_update_properties_cs_d: an output aliases C or C_s");
  int st = covof::vof_corner_fix(h->c, C);
  if (!st) st = covof::vof_properties(h->c, C, rho1, rho2, mu1, mu2, rho, rho_u, rho_v, mu, mu_node, C_s);
  return hip_status(st);
}

extern "C" int covof_ctx_set_obstacle(covof_ctx *h, const covof_obstacle *obs) {
  if (!h) return fail(COVOF_BAD_ARGUMENT, "null context");
  covof::Obstacle o;
  if (obs && obs->enabled) {
    if (obs->type != 1 && obs->type != 2)
      return fail(COVOF_BAD_ARGUMENT, "covof_ctx_set_obstacle: type must be 1 (box) or 2 (cylinder)");
    o.enabled = 1;
    o.type = obs->type;
    o.x0 = obs->x0;
    o.y0 = obs->y0;
    o.x1 = obs->x1;
    o.y1 = obs->y1;
    o.xc = obs->xc;
    o.yc = obs->yc;
    o.r = obs->r;
    o.perm = obs->perm;
  }
  h->c.obs = o;
  ++h->c.ps.assembly_id;  // preconditioner records built from an earlier assembly no longer apply
  return COVOF_OK;
}

extern "C" int covof_ctx_set_vof_top_profile(covof_ctx *h, const char *profile, double jet_center,
                                             double jet_width) {
  if (!h || !profile) return fail(COVOF_BAD_ARGUMENT, "null argument");
  std::string s(profile);
  while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
  if (s == "JET_INLET") {
    h->c.jet.on = 1;
    h->c.jet.center = jet_center;
    h->c.jet.width = jet_width;
  } else if (s == "UNIFORM" || s.empty()) {
    h->c.jet.on = 0;
  } else {
    return fail(COVOF_BAD_ARGUMENT, "covof_ctx_set_vof_top_profile: profile must be UNIFORM or JET_INLET");
  }
  return COVOF_OK;
}

static bool same_halo_array(const double *a, const double *b) { return a && b && a == b; }

extern "C" int covof_curvature_csf_method_d(covof_ctx *h, const double *G, const double *C, int enable,
                                            double sigma, int method, double *curvature, double *fx_sigma,
                                            double *fy_sigma) {
  if (!h || !C || !curvature || !fx_sigma || !fy_sigma) return fail(COVOF_BAD_ARGUMENT, "null argument");
  if (method != COVOF_CURVATURE_NORMALS && method != COVOF_CURVATURE_HEIGHT_FUNCTION)
    return fail(COVOF_BAD_ARGUMENT, "covof_curvature_csf: unknown curvature method");
  if (!G) G = C;
  const double *in[2] = {G, C};
  const double *out[3] = {curvature, fx_sigma, fy_sigma};
  for (const double *o : out)
    for (const double *q : in)
      if (same_halo_array(o, q)) return fail(COVOF_BAD_ARGUMENT, "covof_curvature_csf: an output aliases an input");
  if (curvature == fx_sigma || curvature == fy_sigma || fx_sigma == fy_sigma)
    return fail(COVOF_BAD_ARGUMENT, "covof_curvature_csf: outputs must be distinct");
  return hip_status(covof::curvature_csf(h->c, G, C, enable, sigma, curvature, fx_sigma, fy_sigma, method));
}

extern "C" int covof_curvature_csf_d(covof_ctx *h, const double *G, const double *C, int enable, double sigma,
                                     double *curvature, double *fx_sigma, double *fy_sigma) {
  return covof_curvature_csf_method_d(h, G, C, enable, sigma, COVOF_CURVATURE_NORMALS, curvature, fx_sigma,
                                      fy_sigma);
}

extern "C" int covof_smooth_diffusion_d(covof_ctx *h, const double *C, int steps, double *C_tilde) {
  if (!h || !C || !C_tilde) return fail(COVOF_BAD_ARGUMENT, "null argument");
  if (C == C_tilde) return fail(COVOF_BAD_ARGUMENT, "covof_smooth_diffusion: C_tilde must not alias C");
  return hip_status(covof::smooth_diffusion(h->c, C, steps, C_tilde));
}

extern "C" int covof_svof_smooth_d(covof_ctx *h, const double *C, double filter_width, double *C_s) {
  if (!h || !C || !C_s) return fail(COVOF_BAD_ARGUMENT, "null argument");
  if (C == C_s) return fail(COVOF_BAD_ARGUMENT, "covof_svof_smooth: C_s must not alias C");
  return hip_status(covof::svof_smooth(h->c, C, filter_width, C_s));
}

// Host-array forms: copy in, run, copy out (the outputs' previous contents
// travel too: the force arrays keep their entries outside the face ranges).
extern "C" int covof_curvature_csf(covof_ctx *h, const double *G, const double *C, int enable, double sigma,
                                   double *curvature, double *fx_sigma, double *fy_sigma) {
  if (!h || !C || !curvature || !fx_sigma || !fy_sigma) return fail(COVOF_BAD_ARGUMENT, "null argument");
  double *d[5];
  for (int k = 0; k < 5; ++k)
    if (!(d[k] = mirror(h, k))) return hip_status(COVOF_HIP_ERROR);
  int st = h2d(h, d[0], C);
  if (!st && G && G != C) st = h2d(h, d[1], G);
  if (!st) st = h2d(h, d[3], fx_sigma);
  if (!st) st = h2d(h, d[4], fy_sigma);
  if (st) return hip_status(st);
  st = covof_curvature_csf_d(h, (G && G != C) ? d[1] : d[0], d[0], enable, sigma, d[2], d[3], d[4]);
  if (st) return st;
  st = d2h(h, curvature, d[2]);
  if (!st) st = d2h(h, fx_sigma, d[3]);
  if (!st) st = d2h(h, fy_sigma, d[4]);
  if (!st && hipStreamSynchronize(h->c.stream) != hipSuccess) st = COVOF_HIP_ERROR;
  return hip_status(st);
}

extern "C" int covof_smooth_diffusion(covof_ctx *h, const double *C, int steps, double *C_tilde) {
  if (!h || !C || !C_tilde) return fail(COVOF_BAD_ARGUMENT, "null argument");
  double *d0 = mirror(h, 0), *d1 = mirror(h, 1);
  if (!d0 || !d1) return hip_status(COVOF_HIP_ERROR);
  int st = h2d(h, d0, C);
  if (st) return hip_status(st);
  st = covof_smooth_diffusion_d(h, d0, steps, d1);
  if (st) return st;
  st = d2h(h, C_tilde, d1);
  if (!st && hipStreamSynchronize(h->c.stream) != hipSuccess) st = COVOF_HIP_ERROR;
  return hip_status(st);
}

extern "C" int covof_svof_smooth(covof_ctx *h, const double *C, double filter_width, double *C_s) {
  if (!h || !C || !C_s) return fail(COVOF_BAD_ARGUMENT, "null argument");
  double *d0 = mirror(h, 0), *d1 = mirror(h, 1);
  if (!d0 || !d1) return hip_status(COVOF_HIP_ERROR);
  int st = h2d(h, d0, C);
  if (st) return hip_status(st);
  st = covof_svof_smooth_d(h, d0, filter_width, d1);
  if (st) return st;
  st = d2h(h, C_s, d1);
  if (!st && hipStreamSynchronize(h->c.stream) != hipSuccess) st = COVOF_HIP_ERROR;
  return hip_status(st);
}

extern "C" int covof_poisson_assemble_d(covof_ctx *h, const double *u, const double *v, const double *rho_u,
                                        const double *rho_v, double dt, double *b) {
  if (!h || !u || !v || !rho_u || !rho_v) return fail(COVOF_BAD_ARGUMENT, "null argument");
  return hip_status(covof::poisson_assemble(h->c, u, v, rho_u, rho_v, dt, b));
}

extern "C" int covof_poisson_matvec_d(covof_ctx *h, const double *x, double *y) {
  if (!h || !x || !y) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const int st = covof::poisson_matvec(h->c, x, y);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "operator not assembled");
  return hip_status(st);
}

extern "C" int covof_poisson_precond_d(covof_ctx *h, int precond, const double *r, double *z) {
  if (!h || !r || !z) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const int st = covof::poisson_precond(h->c, precond, r, z);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "operator not assembled or unknown preconditioner");
  return hip_status(st);
}

extern "C" int covof_poisson_solve_d(covof_ctx *h, const double *b, double *x, int lin_solver, int precond,
                                     double tol, int maxit, int precision, int *iters, double *relres) {
  if (!h || !b || !x) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const int st = covof::poisson_solve(h->c, b, x, lin_solver, precond, tol, maxit, precision, iters, relres);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "unsupported solver/preconditioner/precision combination or operator not assembled");
  return hip_status(st);
}

extern "C" int covof_sp_poisson_d(covof_ctx *h, const double *u, const double *v, const double *rho_u,
                                  const double *rho_v, double *p, double dt, int lin_solver, int precond,
                                  double tol, int maxit, int precision, int *iters, double *relres) {
  if (!h || !u || !v || !rho_u || !rho_v || !p) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const int st = covof::sp_poisson(h->c, u, v, rho_u, rho_v, p, dt, lin_solver, precond, tol, maxit,
                                   precision, iters, relres);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "unsupported solver/preconditioner/precision combination");
  return hip_status(st);
}

extern "C" int covof_sp_poisson(covof_ctx *h, const double *u, const double *v, const double *rho_u,
                                const double *rho_v, double *p, double dt, int lin_solver, int precond,
                                double tol, int maxit, int precision, int *iters, double *relres) {
  if (!h || !u || !v || !rho_u || !rho_v || !p) return fail(COVOF_BAD_ARGUMENT, "null argument");
  double *d[5];
  for (int k = 0; k < 5; ++k)
    if (!(d[k] = mirror(h, k))) return hip_status(COVOF_HIP_ERROR);
  const double *hs[5] = {u, v, rho_u, rho_v, p};
  int st = 0;
  for (int k = 0; k < 5 && !st; ++k) st = h2d(h, d[k], hs[k]);
  if (st) return hip_status(st);
  const int sst = covof_sp_poisson_d(h, d[0], d[1], d[2], d[3], d[4], dt, lin_solver, precond, tol, maxit,
                                     precision, iters, relres);
  if (sst == COVOF_BAD_ARGUMENT || sst == COVOF_HIP_ERROR) return sst;
  st = d2h(h, p, d[4]);
  if (!st && hipStreamSynchronize(h->c.stream) != hipSuccess) st = COVOF_HIP_ERROR;
  return st ? hip_status(st) : sst;
}

extern "C" int covof_sp_correct_d(covof_ctx *h, const double *p, const double *rho_u, const double *rho_v,
                                  double dt, double *u, double *v) {
  if (!h || !p || !rho_u || !rho_v || !u || !v) return fail(COVOF_BAD_ARGUMENT, "null argument");
  return hip_status(covof::sp_correct(h->c, p, rho_u, rho_v, dt, u, v));
}

extern "C" int covof_sp_correct(covof_ctx *h, const double *p, const double *rho_u, const double *rho_v,
                                double dt, double *u, double *v) {
  if (!h || !p || !rho_u || !rho_v || !u || !v) return fail(COVOF_BAD_ARGUMENT, "null argument");
  double *d[5];
  for (int k = 0; k < 5; ++k)
    if (!(d[k] = mirror(h, k))) return hip_status(COVOF_HIP_ERROR);
  const double *hs[5] = {p, rho_u, rho_v, u, v};
  int st = 0;
  for (int k = 0; k < 5 && !st; ++k) st = h2d(h, d[k], hs[k]);
  if (!st) st = covof::sp_correct(h->c, d[0], d[1], d[2], dt, d[3], d[4]);
  if (!st) st = d2h(h, u, d[3]);
  if (!st) st = d2h(h, v, d[4]);
  if (!st && hipStreamSynchronize(h->c.stream) != hipSuccess) st = COVOF_HIP_ERROR;
  return hip_status(st);
}

extern "C" int covof_poisson_adjoint_d(covof_ctx *h, const double *rho, double dt, const double *x,
                                       const double *w, double *dJ_db, double *dJ_dlogrho, int lin_solver,
                                       int precond, double tol, int maxit, int precision, int *iters,
                                       double *relres) {
  if (!h || !rho || !x || !w || !dJ_db || !dJ_dlogrho) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const int st = covof::poisson_adjoint(h->c, rho, dt, x, w, dJ_db, dJ_dlogrho, lin_solver, precond, tol,
                                        maxit, precision, iters, relres);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "unsupported solver combination");
  return hip_status(st);
}

extern "C" int covof_sp_pressure_d(covof_ctx *h, const double *x, double *p) {
  if (!h || !x || !p) return fail(COVOF_BAD_ARGUMENT, "null argument");
  return hip_status(covof::sp_pressure(h->c, x, p));
}

extern "C" int covof_projection_adjoint_d(covof_ctx *h, const double *rho, const double *rho_u,
                                          const double *rho_v, double dt, const double *x, const double *w_u,
                                          const double *w_v, const double *w_p, double *dJ_du, double *dJ_dv,
                                          double *dJ_dlogrho, int precond, double tol, int maxit, int precision,
                                          int *iters, double *relres) {
  if (!h || !rho || !rho_u || !rho_v || !x || !w_u || !w_v || !dJ_du || !dJ_dv || !dJ_dlogrho)
    return fail(COVOF_BAD_ARGUMENT, "null argument");
  if (dJ_du == w_u || dJ_dv == w_v || dJ_du == dJ_dv)
    return fail(COVOF_BAD_ARGUMENT, "covof_projection_adjoint_d: outputs must not alias the objective's gradients");
  const int st = covof::projection_adjoint(h->c, rho, rho_u, rho_v, dt, x, w_u, w_v, w_p, dJ_du, dJ_dv, dJ_dlogrho,
                                           precond, tol, maxit, precision, iters, relres);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "operator not assembled or unsupported preconditioner/precision");
  return hip_status(st);
}

// ------------------------------------------- multi-step adjoint of the kinetic energy
extern "C" int covof_vof_advect_adjoint_d(covof_ctx *h, const double *C, const double *u, const double *v,
                                          double dt, const char *reconstruction_method, int step_count,
                                          int two_phase, int fold_input_halo, const double *dJ_dCnew,
                                          double *dJ_dC, double *dJ_du, double *dJ_dv) {
  int code;
  if (!h || !C || !u || !v || !dJ_dCnew || !dJ_dC || !dJ_du || !dJ_dv)
    return fail(COVOF_BAD_ARGUMENT, "null argument");
  if (recon_code(reconstruction_method, &code))
    return fail(COVOF_BAD_ARGUMENT, "reconstruction method must be PLIC or ELVIRA");
  if (!(dt >= 0.0) || step_count < 0) return fail(COVOF_BAD_ARGUMENT, "dt and step_count must be >= 0");
  if (dJ_du == dJ_dv || dJ_dC == dJ_du || dJ_dC == dJ_dv || dJ_dCnew == dJ_du || dJ_dCnew == dJ_dv ||
      C == dJ_dC || u == dJ_du || v == dJ_dv)
    return fail(COVOF_BAD_ARGUMENT, "covof_vof_advect_adjoint_d: outputs must not alias other arguments");
  const int st = covof::vof_advect_adjoint(h->c, code, C, u, v, dt, step_count, two_phase, fold_input_halo,
                                           dJ_dCnew, dJ_dC, dJ_du, dJ_dv);
  if (st == COVOF_BAD_ARGUMENT)
    return fail(st, "covof_vof_advect_adjoint_d: the mesh must be periodic in both directions or in neither, "
                    "with the UNIFORM top profile");
  return hip_status(st);
}

extern "C" int covof_vof_fill_halo_d(covof_ctx *h, double *C) {
  if (!h || !C) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const int st = covof::vof_fill_halo(h->c, C);
  if (st == COVOF_BAD_ARGUMENT)
    return fail(st, "covof_vof_fill_halo_d: the mesh must be periodic in both directions or in neither");
  return hip_status(st);
}

extern "C" int covof_kinetic_energy_d(covof_ctx *h, const double *C, const double *u, const double *v, double rho1,
                                      double rho2, double *ke, double *dKE_dC, double *dKE_du, double *dKE_dv) {
  if (!h || !C || !u || !v) return fail(COVOF_BAD_ARGUMENT, "null argument");
  if ((dKE_dC && (dKE_dC == C || dKE_dC == u || dKE_dC == v)) || (dKE_du && (dKE_du == u || dKE_du == C)) ||
      (dKE_dv && (dKE_dv == v || dKE_dv == C)) || (dKE_du && dKE_du == dKE_dv))
    return fail(COVOF_BAD_ARGUMENT, "covof_kinetic_energy_d: outputs must not alias the inputs or each other");
  return hip_status(covof::kinetic_energy(h->c, C, u, v, rho1, rho2, ke, dKE_dC, dKE_du, dKE_dv));
}

extern "C" int covof_properties_adjoint_d(covof_ctx *h, const double *C, double rho1, double rho2,
                                          const double *dJ_dlogrho, double *dJ_dC) {
  if (!h || !C || !dJ_dlogrho || !dJ_dC) return fail(COVOF_BAD_ARGUMENT, "null argument");
  if (dJ_dC == C) return fail(COVOF_BAD_ARGUMENT, "covof_properties_adjoint_d: dJ_dC must not alias C");
  return hip_status(covof::properties_adjoint(h->c, C, rho1, rho2, dJ_dlogrho, dJ_dC));
}

extern "C" int covof_inertial_predictor_d(covof_ctx *h, int first, double dt, double gx, double gy, const double *u,
                                          const double *v, const double *u_nm1, const double *v_nm1,
                                          double *u_star, double *v_star) {
  if (!h || !u || !v || !u_star || !v_star || (!first && (!u_nm1 || !v_nm1)))
    return fail(COVOF_BAD_ARGUMENT, "null argument");
  if (u_star == v_star || u_star == u || v_star == v || u_star == u_nm1 || v_star == v_nm1)
    return fail(COVOF_BAD_ARGUMENT, "covof_inertial_predictor_d: outputs must not alias the inputs");
  const int st = covof::inertial_predictor(h->c, first, dt, gx, gy, u, v, u_nm1, v_nm1, u_star, v_star);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "every side must be a wall (DIRICHLET, FREESLIP) or periodic");
  return hip_status(st);
}

extern "C" int covof_inertial_predictor_adjoint_d(covof_ctx *h, int first, const double *w_u, const double *w_v,
                                                  double *dJ_du, double *dJ_dv, double *dJ_du_nm1,
                                                  double *dJ_dv_nm1) {
  if (!h || !w_u || !w_v || !dJ_du || !dJ_dv || (!first && (!dJ_du_nm1 || !dJ_dv_nm1)))
    return fail(COVOF_BAD_ARGUMENT, "null argument");
  if (dJ_du == dJ_dv || dJ_du == w_u || dJ_dv == w_v || (!first && (dJ_du_nm1 == dJ_du || dJ_dv_nm1 == dJ_dv ||
                                                                     dJ_du_nm1 == dJ_dv_nm1)))
    return fail(COVOF_BAD_ARGUMENT, "covof_inertial_predictor_adjoint_d: outputs must be distinct arrays");
  const int st = covof::inertial_predictor_adjoint(h->c, first, w_u, w_v, dJ_du, dJ_dv, dJ_du_nm1, dJ_dv_nm1);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "every side must be a wall (DIRICHLET, FREESLIP) or periodic");
  return hip_status(st);
}

// ------------------------------------------------------------- momentum block
static int mom_cfg(const covof_momentum_cfg *in, covof::MomCfg *out) {
  if (!in) return COVOF_BAD_ARGUMENT;
  if (in->time_scheme < 1 || in->time_scheme > 2 || in->convective_scheme < 1 || in->convective_scheme > 3 ||
      in->convective_linearization < 1 || in->convective_linearization > 3 ||
      (in->solver_type != 0 && in->solver_type != COVOF_SOLVER_MONOLITHIC &&
       in->solver_type != COVOF_SOLVER_STANDARD_PROJECTION))
    return COVOF_BAD_ARGUMENT;
  out->gear = in->time_scheme == 2 && in->has_nm1;
  out->conv = in->enable_convection != 0;
  out->scheme = in->convective_scheme;
  out->lin = in->convective_linearization;
  out->ab_nm1 = in->convective_linearization == 2 && in->has_nm1;
  out->gx = in->gx;
  out->gy = in->gy;
  for (int k = 0; k < 4; ++k) {
    out->bcu[k] = in->bc_u[k];
    out->bcv[k] = in->bc_v[k];
  }
  out->mono = in->solver_type == COVOF_SOLVER_MONOLITHIC;
  return COVOF_OK;
}

extern "C" int covof_momentum_size(covof_ctx *h, int solver_type, int *n, int *u_nx, int *v_ny) {
  if (!h) return fail(COVOF_BAD_ARGUMENT, "null context");
  const covof::Mesh &m = h->c.m;
  const int a = m.nx + (m.perx ? 0 : 1), b = m.ny + (m.pery ? 0 : 1);
  if (u_nx) *u_nx = a;
  if (v_ny) *v_ny = b;
  if (n) *n = a * m.ny + m.nx * b + (solver_type == COVOF_SOLVER_MONOLITHIC ? m.nx * m.ny : 0);
  return COVOF_OK;
}

extern "C" int covof_momentum_set_fields_d(covof_ctx *h, const covof_momentum_cfg *cfg, const double *u,
                                           const double *v, const double *u_old, const double *v_old,
                                           const double *u_nm1, const double *v_nm1, const double *rho_u,
                                           const double *rho_v, const double *mu, const double *mu_node,
                                           const double *fx_sigma, const double *fy_sigma, double dt) {
  if (!h) return fail(COVOF_BAD_ARGUMENT, "null context");
  covof::MomCfg c;
  if (mom_cfg(cfg, &c)) return fail(COVOF_BAD_ARGUMENT, "covof_momentum_cfg: selector out of range");
  const covof::MomFields f{u, v, u_old, v_old, u_nm1, v_nm1, rho_u, rho_v, mu, mu_node, fx_sigma, fy_sigma};
  const int st = covof::mom_bind(h->c, c, f, dt);
  if (st == COVOF_BAD_ARGUMENT)
    return fail(st, "momentum block: null field, dt <= 0, missing n-1 fields, unsupported boundary sides "
                    "(periodic sides in pairs, no NEUMANN top), or an obstacle set on the context");
  return hip_status(st);
}

extern "C" int covof_momentum_diag_d(covof_ctx *h, double *diag) {
  if (!h || !diag) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const int st = covof::mom_diag(h->c, diag);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "momentum fields not set");
  return hip_status(st);
}

extern "C" int covof_momentum_rhs_d(covof_ctx *h, double *rhs) {
  if (!h || !rhs) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const int st = covof::mom_rhs(h->c, rhs);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "momentum fields not set");
  return hip_status(st);
}

extern "C" int covof_momentum_matvec_d(covof_ctx *h, const double *x, double *y) {
  if (!h || !x || !y) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const int st = covof::mom_matvec(h->c, x, y);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "momentum fields not set");
  return hip_status(st);
}

extern "C" int covof_momentum_solve_d(covof_ctx *h, const double *b, double *x, int lin_solver, int precond,
                                      double tol, int maxit, int *iters, double *relres) {
  if (!h || !b || !x) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const int st = covof::mom_solve(h->c, b, x, lin_solver, precond, tol, maxit, iters, relres);
  if (st == COVOF_BAD_ARGUMENT)
    return fail(st, "momentum fields not set, or solver not BICGSTAB2/CG, or preconditioner not "
                    "NONE/JACOBI/SSOR/BLOCK_TRIANGULAR");
  return hip_status(st);
}

extern "C" int covof_momentum_precond_d(covof_ctx *h, int precond, const double *r, double *z) {
  if (!h || !r || !z) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const int st = covof::mom_apply_precond(h->c, precond, r, z);
  if (st == COVOF_BAD_ARGUMENT)
    return fail(st, "momentum fields not set, or preconditioner not NONE/JACOBI/SSOR/BLOCK_TRIANGULAR");
  return hip_status(st);
}

// Host-array forms for a caller whose fields live on the host (the
// PRECOND_CUSTOM hook of a CPU solve): the fields go to the context's device
// mirrors 0-11, which stay bound until the next host entry point reuses them.
extern "C" int covof_momentum_set_fields(covof_ctx *h, const covof_momentum_cfg *cfg, const double *u,
                                         const double *v, const double *u_old, const double *v_old,
                                         const double *u_nm1, const double *v_nm1, const double *rho_u,
                                         const double *rho_v, const double *mu, const double *mu_node,
                                         const double *fx_sigma, const double *fy_sigma, double dt) {
  if (!h || !u || !v || !u_old || !v_old || !rho_u || !rho_v || !mu || !mu_node || !fx_sigma || !fy_sigma)
    return fail(COVOF_BAD_ARGUMENT, "null argument");
  const double *hs[12] = {u, v, u_old, v_old, u_nm1, v_nm1, rho_u, rho_v, mu, mu_node, fx_sigma, fy_sigma};
  double *d[12] = {nullptr};
  int st = 0;
  for (int k = 0; k < 12 && !st; ++k) {
    if (!hs[k]) continue;
    if (!(d[k] = mirror(h, k))) return hip_status(COVOF_HIP_ERROR);
    st = h2d(h, d[k], hs[k]);
  }
  if (st) return hip_status(st);
  return covof_momentum_set_fields_d(h, cfg, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10],
                                     d[11], dt);
}

extern "C" int covof_momentum_precond(covof_ctx *h, int precond, int n, const double *r, double *z) {
  if (!h || !r || !z) return fail(COVOF_BAD_ARGUMENT, "null argument");
  const Ctx &c = h->c;
  if (!c.mom.bound) return fail(COVOF_BAD_ARGUMENT, "momentum fields not set");
  if (n != c.mom.n) return fail(COVOF_BAD_ARGUMENT, "covof_momentum_precond: n is not the bound system's size");
  int nmax = 0;
  covof_momentum_size(h, COVOF_SOLVER_MONOLITHIC, &nmax, nullptr, nullptr);
  for (double *&p : h->hvec)
    if (!p && hipMalloc(&p, (size_t)nmax * sizeof(double)) != hipSuccess) return hip_status(COVOF_HIP_ERROR);
  const size_t bytes = (size_t)n * sizeof(double);
  if (hipMemcpyAsync(h->hvec[0], r, bytes, hipMemcpyHostToDevice, c.stream) != hipSuccess)
    return hip_status(COVOF_HIP_ERROR);
  const int st = covof_momentum_precond_d(h, precond, h->hvec[0], h->hvec[1]);
  if (st) return st;
  if (hipMemcpyAsync(z, h->hvec[1], bytes, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
      hipStreamSynchronize(c.stream) != hipSuccess)
    return hip_status(COVOF_HIP_ERROR);
  return COVOF_OK;
}

extern "C" int covof_velocity_bc_d(covof_ctx *h, const covof_momentum_cfg *cfg, double *u, double *v,
                                   int homogeneous) {
  if (!h || !u || !v) return fail(COVOF_BAD_ARGUMENT, "null argument");
  covof::MomCfg c;
  if (mom_cfg(cfg, &c)) return fail(COVOF_BAD_ARGUMENT, "covof_momentum_cfg: selector out of range");
  const int st = covof::velocity_bc(h->c, c, u, v, homogeneous);
  if (st == COVOF_BAD_ARGUMENT) return fail(st, "periodic sides must come in pairs");
  return hip_status(st);
}

extern "C" int covof_sp_predictor_d(covof_ctx *h, const covof_momentum_cfg *cfg, double *u, double *v,
                                    const double *u_old, const double *v_old, const double *u_nm1,
                                    const double *v_nm1, const double *rho_u, const double *rho_v,
                                    const double *mu, const double *mu_node, const double *fx_sigma,
                                    const double *fy_sigma, double dt, int lin_solver, int precond, double tol,
                                    int maxit, int *iters, double *relres) {
  int st = covof_momentum_set_fields_d(h, cfg, u, v, u_old, v_old, u_nm1, v_nm1, rho_u, rho_v, mu, mu_node,
                                       fx_sigma, fy_sigma, dt);
  if (st) return st;
  st = covof::mom_predictor(h->c, u, v, lin_solver, precond, tol, maxit, iters, relres);
  if (st == COVOF_BAD_ARGUMENT)
    return fail(st, "monolithic cfg, solver not BICGSTAB2/CG, or preconditioner not NONE/JACOBI/SSOR");
  return hip_status(st);
}

extern "C" int covof_sp_predictor(covof_ctx *h, const covof_momentum_cfg *cfg, double *u, double *v,
                                  const double *u_old, const double *v_old, const double *u_nm1,
                                  const double *v_nm1, const double *rho_u, const double *rho_v,
                                  const double *mu, const double *mu_node, const double *fx_sigma,
                                  const double *fy_sigma, double dt, int lin_solver, int precond, double tol,
                                  int maxit, int *iters, double *relres) {
  if (!h || !u || !v || !u_old || !v_old || !rho_u || !rho_v || !mu || !mu_node || !fx_sigma || !fy_sigma)
    return fail(COVOF_BAD_ARGUMENT, "null argument");
  const double *hs[12] = {u, v, u_old, v_old, u_nm1, v_nm1, rho_u, rho_v, mu, mu_node, fx_sigma, fy_sigma};
  double *d[12] = {nullptr};
  int st = 0;
  for (int k = 0; k < 12 && !st; ++k) {
    if (!hs[k]) continue;
    if (!(d[k] = mirror(h, k))) return hip_status(COVOF_HIP_ERROR);
    st = h2d(h, d[k], hs[k]);
  }
  if (st) return hip_status(st);
  const int sst = covof_sp_predictor_d(h, cfg, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10],
                                       d[11], dt, lin_solver, precond, tol, maxit, iters, relres);
  if (sst == COVOF_BAD_ARGUMENT || sst == COVOF_HIP_ERROR) return sst;
  st = d2h(h, u, d[0]);
  if (!st) st = d2h(h, v, d[1]);
  if (!st && hipStreamSynchronize(h->c.stream) != hipSuccess) st = COVOF_HIP_ERROR;
  return st ? hip_status(st) : sst;
}

extern "C" int covof_monolithic_step_d(covof_ctx *h, const covof_momentum_cfg *cfg, double *u, double *v,
                                      double *p, const double *u_old, const double *v_old, const double *u_nm1,
                                      const double *v_nm1, const double *rho_u, const double *rho_v,
                                      const double *mu, const double *mu_node, const double *fx_sigma,
                                      const double *fy_sigma, double dt, int lin_solver, int precond, double tol,
                                      int maxit, int *iters, double *relres) {
  if (!cfg || cfg->solver_type != COVOF_SOLVER_MONOLITHIC)
    return fail(COVOF_BAD_ARGUMENT, "covof_monolithic_step_d: cfg->solver_type must be SOLVER_MONOLITHIC");
  if (!p) return fail(COVOF_BAD_ARGUMENT, "null argument");
  int st = covof_momentum_set_fields_d(h, cfg, u, v, u_old, v_old, u_nm1, v_nm1, rho_u, rho_v, mu, mu_node,
                                       fx_sigma, fy_sigma, dt);
  if (st) return st;
  st = covof::mono_step(h->c, u, v, p, lin_solver, precond, tol, maxit, iters, relres);
  if (st == COVOF_BAD_ARGUMENT)
    return fail(st, "solver not BICGSTAB2/CG, or preconditioner not NONE/JACOBI/SSOR/BLOCK_TRIANGULAR");
  return hip_status(st);
}

extern "C" int covof_monolithic_step(covof_ctx *h, const covof_momentum_cfg *cfg, double *u, double *v, double *p,
                                     const double *u_old, const double *v_old, const double *u_nm1,
                                     const double *v_nm1, const double *rho_u, const double *rho_v,
                                     const double *mu, const double *mu_node, const double *fx_sigma,
                                     const double *fy_sigma, double dt, int lin_solver, int precond, double tol,
                                     int maxit, int *iters, double *relres) {
  if (!h || !u || !v || !p || !u_old || !v_old || !rho_u || !rho_v || !mu || !mu_node || !fx_sigma || !fy_sigma)
    return fail(COVOF_BAD_ARGUMENT, "null argument");
  const double *hs[13] = {u, v, u_old, v_old, u_nm1, v_nm1, rho_u, rho_v, mu, mu_node, fx_sigma, fy_sigma, p};
  double *d[13] = {nullptr};
  int st = 0;
  for (int k = 0; k < 13 && !st; ++k) {
    if (!hs[k]) continue;
    if (!(d[k] = mirror(h, k))) return hip_status(COVOF_HIP_ERROR);
    st = h2d(h, d[k], hs[k]);
  }
  if (st) return hip_status(st);
  const int sst = covof_monolithic_step_d(h, cfg, d[0], d[1], d[12], d[2], d[3], d[4], d[5], d[6], d[7], d[8],
                                          d[9], d[10], d[11], dt, lin_solver, precond, tol, maxit, iters, relres);
  if (sst == COVOF_BAD_ARGUMENT || sst == COVOF_HIP_ERROR) return sst;
  st = d2h(h, u, d[0]);
  if (!st) st = d2h(h, v, d[1]);
  if (!st) st = d2h(h, p, d[12]);
  if (!st && hipStreamSynchronize(h->c.stream) != hipSuccess) st = COVOF_HIP_ERROR;
  return st ? hip_status(st) : sst;
}

// ------------------------------------------------------------------ output
extern "C" int covof_history_d(covof_ctx *h, const double *C, const double *u, const double *v, const double *p,
                               const double *rho, const double *T, double *out) {
  if (!h || !C || !u || !v || !p || !rho || !out) return fail(COVOF_BAD_ARGUMENT, "null argument");
  return hip_status(covof::history_row(h->c, C, u, v, p, rho, T, out));
}

extern "C" int covof_cell_output_d(covof_ctx *h, const double *u, const double *v, double *div, double *uc,
                                   double *vc) {
  if (!h || !u || !v || !div || !uc || !vc) return fail(COVOF_BAD_ARGUMENT, "null argument");
  return hip_status(covof::cell_output(h->c, u, v, div, uc, vc));
}

// ------------------------------------------------------------------ profiling
namespace covof {
void prof_events(Ctx &c, int tag, hipEvent_t *start, hipEvent_t *stop) {
  Prof &P = c.prof;
  *start = *stop = nullptr;
  if (!P.on || tag < 0 || tag >= COVOF_NUM_KERNEL_TAGS) return;
  if (!P.created) {
    for (int k = 0; k < 2 * 1024; ++k)
      if (hipEventCreate(&P.ev[k]) != hipSuccess) {
        P.on = 0;
        return;
      }
    P.created = 1;
  }
  if (P.used >= 1024) {  // pool full: wait for the stream and fold the records
    if (hipStreamSynchronize(c.stream) != hipSuccess) return;
    prof_harvest(c);
  }
  P.tag[P.used] = tag;
  *start = P.ev[2 * P.used];
  *stop = P.ev[2 * P.used + 1];
  ++P.used;
}
void prof_harvest(Ctx &c) {
  Prof &P = c.prof;
  for (int k = 0; k < P.used; ++k) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, P.ev[2 * k], P.ev[2 * k + 1]) == hipSuccess) {
      P.total_ms[P.tag[k]] += ms;
      P.count[P.tag[k]] += 1;
    }
  }
  P.used = 0;
}
}  // namespace covof

extern "C" int covof_ctx_profile(covof_ctx *h, int on) {
  if (!h) return fail(COVOF_BAD_ARGUMENT, "null context");
  covof::Prof &P = h->c.prof;
  if (hipStreamSynchronize(h->c.stream) != hipSuccess) return hip_status(COVOF_HIP_ERROR);
  covof::prof_harvest(h->c);
  if (on) {
    for (int t = 0; t < COVOF_NUM_KERNEL_TAGS; ++t) {
      P.total_ms[t] = 0.0;
      P.count[t] = 0;
    }
  }
  P.on = on > 0 ? on : 0;
  return COVOF_OK;
}

extern "C" int covof_ctx_profile_read(covof_ctx *h, int tag, double *total_ms, long *count) {
  if (!h || tag < 0 || tag >= COVOF_NUM_KERNEL_TAGS || !total_ms || !count)
    return fail(COVOF_BAD_ARGUMENT, "covof_ctx_profile_read: bad argument");
  if (hipStreamSynchronize(h->c.stream) != hipSuccess) return hip_status(COVOF_HIP_ERROR);
  covof::prof_harvest(h->c);
  *total_ms = h->c.prof.total_ms[tag];
  *count = h->c.prof.count[tag];
  return COVOF_OK;
}
