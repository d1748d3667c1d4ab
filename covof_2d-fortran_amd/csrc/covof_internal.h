// Internal declarations shared by the HIP translation units of libcovof_hip.
// Not part of the public C-ABI (that is include/covof_hip.h).
#pragma once

#include <cmath>
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>

#include "covof_hip.h"

namespace covof {

enum { BC_DIRICHLET = 1, BC_NEUMANN = 2, BC_PERIODIC = 3, BC_FREESLIP = 4 };

// Geometry of the single-rank mesh (src/module_mesh.f90:67-144).
struct Mesh {
  int nx, ny;
  double lx, ly, dx, dy;
  int bc[4];     // left, right, bottom, top (reference integer BC types)
  int perx, pery;  // Cartesian periodicity (src/module_simulation.f90:115-117)
};

// Device state of one context: every buffer is allocated once at create.
struct Vof {
  double *Cb = nullptr;                 // second C buffer (ping-pong), halo layout
  double *snx = nullptr, *sny = nullptr, *sal = nullptr;  // segments, halo layout, mixed cells only
  uint8_t *cd = nullptr;                // dilation indicator [C^n >= 0.5], interior, halo layout
  int *list[1] = {nullptr};             // per-tile lists of mixed cells (linear halo index)
  int *count = nullptr;                 // per-tile list sizes
  int *offs = nullptr;                  // exclusive prefix of count (ntiles + 1 entries)
  double *adj = nullptr;                // work arrays of the step adjoint (made on its first call)
};

// Work arrays of the surface steps (surface.hip), made on first use: two
// halo-size FP64 arrays.
struct Surface {
  double *work = nullptr;
  // the SVOF smoothing on the strip sweeps: records, two state buffers, halo rows
  double *sv_rec = nullptr, *sv_xa = nullptr, *sv_xb = nullptr, *sv_h = nullptr;
};

// Options of the SP momentum block (momentum.hip), the covof_momentum_cfg of
// the C-ABI with the derived switches.
struct MomCfg {
  int gear;      // time_scheme == GEAR with a valid n-1 state: alpha_time 1.5 and BDF2 inertia
  int conv;      // enable_convection
  int scheme;    // 1 UPWIND, 2 CENTERED, other: TVD superbee
  int lin;       // 1 PICARD, 2 ADAMS_BASHFORTH, 3 NEWTON
  int ab_nm1;    // ADAMS_BASHFORTH with a valid n-1 state: extrapolated advecting velocity
  double gx, gy;
  double bcu[4], bcv[4];
  int mono;      // SOLVER_MONOLITHIC: unknowns u, v, p; pressure gradient and continuity rows
};
// The fields the momentum operator, its diagonal and right-hand side read
// (the reference's fields_old), halo layout.
struct MomFields {
  const double *u, *v, *uo, *vo, *un, *vn, *ru, *rv, *mu, *mn, *fx, *fy;
};
// Device state of the momentum block, made on first use.
struct Mom {
  int bound = 0;              // fields and options set
  int u_nx = 0, v_ny = 0;     // unknowns per row of u, rows of v
  long nu = 0, nuv = 0, n = 0;  // u unknowns, u + v unknowns, all unknowns (+ p when monolithic)
  MomCfg cfg;
  MomFields f;
  double dt = 0.0;
  double *tu = nullptr, *tv = nullptr, *tp = nullptr;  // the operator's persistent work fields (halo layout, zero at first use)
  double *diag = nullptr, *rhs = nullptr, *x = nullptr;
  double *w[9] = {nullptr};   // Krylov work vectors
  // assemble_preconditioner_blocks' SSOR and Schur coefficients (made on first use):
  // u_a{W,E,S,N,P}, v_a{W,E,S,N,P}, schur_a{W,E,S,N,P}
  double *bk[15] = {nullptr};
  double *aP2[2] = {nullptr};  // u_aP, v_aP on the halo grid (the Schur coefficients' source)
  double *sx[3] = {nullptr};   // sx[2]: the block-triangular pressure correction (halo layout)
  // per block (u, v, Schur), in the sequential sweeps' skewed order: the
  // coefficient + rhs records and the two state buffers; and the four halo
  // snapshots (W, E, S, N)
  double *sk[3][3] = {{nullptr}};
  double *hal[3][4] = {{nullptr}};
};

// A linear operator with its preconditioner for the generic Krylov loops
// (krylov.hip); vectors of n entries, device pointers.
struct LinOp {
  long n = 0;
  virtual int apply(const double *x, double *y) = 0;
  virtual int precond(const double *r, double *z) = 0;
  // y = M^-1 (A x); tmp: a work vector the default form may use
  virtual int apply_pa(const double *x, double *y, double *tmp) {
    int st = apply(x, tmp);
    return st ? st : precond(tmp, y);
  }
  virtual ~LinOp() {}
};

// One instantiated capture of a solver cycle (see poisson.hip).
struct PGraph {
  hipGraphExec_t exec = nullptr;
  int key = -1;
};

// Multigrid hierarchy of one precision (multigrid.hip).  Level 0 is the
// assembled operator; levels 1.. own their face coefficients, diagonal,
// correction and right-hand side.
constexpr int MG_MAX_LEVELS = 16;
enum { MG_KIND_VCYCLE = 0, MG_KIND_SGS_RB = 1 };
struct MgLevelBuf {
  int nx = 0, ny = 0;
  void *kx = nullptr, *ky = nullptr, *diag = nullptr, *x = nullptr, *b = nullptr;
};
struct Mg {
  int nlev = 0;
  int built = 0;  // assembly_id the coarse operators were made from
  MgLevelBuf lev[MG_MAX_LEVELS];
  double *rzt = nullptr;  // r.z partials of the finest level's post-pass tiles
};

struct Poisson {
  // face coefficients, FP64: kx((nx+1) x ny, row pitch padded to 64) kx(i,j) = a_W(i,j), kx(nx+1,j) = a_E(nx,j)
  //                          ky(nx x (ny+1))  ky(i,j) = a_S(i,j), ky(i,ny+1) = a_N(i,ny)
  double *kx = nullptr, *ky = nullptr, *diag = nullptr;
  float *kx32 = nullptr, *ky32 = nullptr, *diag32 = nullptr;
  double *b = nullptr, *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr;
  double *t0 = nullptr, *t1 = nullptr, *t2 = nullptr, *t3 = nullptr, *t4 = nullptr, *t5 = nullptr;
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr, *z32 = nullptr;
  float *cg2 = nullptr;     // single-reduction loop: r, w, p, s, two buffers each (8 n)
  double *fx = nullptr, *fy = nullptr;  // per-face adjoint sensitivities (kx / ky layout)
  double *part = nullptr;   // per-block partial sums (3 buffers of nblocks x 4)
  double *scal = nullptr;   // device scalars (see poisson.hip)
  unsigned *ticket = nullptr;  // last-block-done counter
  double *hscal = nullptr;  // pinned host mirror of scal
  int assembled = 0;
  int assembly_id = 0;      // incremented by every assembly
  // the last mixed-precision solve: iterations of the FP32 recurrence and of
  // the FP64 continuation (0 when it did not run)
  int last_it32 = 0, last_it64 = 0;
  // set while an FP64 CG runs with the multigrid cycle in FP32 (the MIXED
  // multigrid solve's default form)
  int mg_fp32_cycle = 0;
  Mg mg64, mg32;
  int nblocks = 0;          // reduction grid size
  PGraph graph[2];          // FP64 cycle, mixed cycle
  hipEvent_t cycle_ev[2] = {nullptr, nullptr};
  // the exact symmetric Gauss-Seidel (PRECOND_SSOR) in the strip sweeps'
  // skewed order: coefficient records, two state buffers, four halo rows;
  // records of the assembly sgs_id
  double *sgs_rec = nullptr, *sgs_xa = nullptr, *sgs_xb = nullptr, *sgs_h = nullptr;
  int sgs_id = -1;
};

// Optional per-launch timing of the hot kernels with HIP events on the
// context's stream (covof_ctx_profile): begin/end events around each launch,
// read back whenever the solvers synchronise anyway.
struct Prof {
  int on = 0;      // sampling period: every on-th iteration of a solve is timed
  hipEvent_t ev[2 * 1024];
  int tag[1024];
  int used = 0;
  int created = 0;
  double total_ms[COVOF_NUM_KERNEL_TAGS] = {0};
  long count[COVOF_NUM_KERNEL_TAGS] = {0};
};

// A solid obstacle of config_t (obs_enabled, obs_type, box, disc, obs_perm),
// for the Poisson assembly and the velocity correction.
struct Obstacle {
  int enabled = 0, type = 0;
  double x0 = 0, y0 = 0, x1 = 0, y1 = 0, xc = 0, yc = 0, r = 0, perm = 1.0e-40;
};

// The top side's VOF profile of apply_vof_bc (cfg%bc_top%profile): JET_INLET
// fills the top halo row with C = 0 inside the jet and 1 outside.
struct TopJet {
  int on = 0;
  double center = 0.5, width = 0.08;
};

// Solver tuning of a context (covof_ctx_set_solver_option); zeros in the
// mixed-mode fields mean "the default of the preconditioner".
// Whether h is a power of two (its reciprocal is then exact, and a quotient
// by h equals the product with that reciprocal bit for bit).
inline bool is_pow2(double h) {
  int e;
  return h > 0.0 && std::frexp(h, &e) == 0.5;
}

struct SolverOpts {
  int mixed_stall = 0;   // reliable updates without a 10 % gain before the FP64 continuation
  int mixed_cycle = 0;   // FP32 iterations per reliable update
  int mixed_loop = 0;    // 1: single-reduction Chronopoulos-Gear loop (Jacobi)
  int mg_coarse = 0;     // 1: series coarse faces (else averaged)
  int mg_sweeps = 1;     // red-black sweeps before and after the coarse correction
  double mg_omega = 1.8; // weight of the coarse-grid correction
  int mg_fused = 1;      // one sweep: fused level passes (k_mg_pre / k_mg_post)
  int mixed_mg = 0;      // MIXED with MG: 0 FP64 CG + FP32 V-cycle, 1 FP32 Krylov + reliable updates
  int mg_project = 1;    // singular operator: mean-free coarsest solve (P S P)
};

struct Ctx {
  Mesh m;
  SolverOpts opts;
  Obstacle obs;
  TopJet jet;
  hipStream_t stream = nullptr;
  int own_stream = 0;
  // COVOF_REDUCTION_REFERENCE: the FP64 Krylov inner products and the
  // pressure mean are serial sums in the reference's element order
  int ref_order = 0;
  Vof vof;
  Surface surf;
  Poisson ps;
  Mom mom;
  Prof prof;
};

// The event pair that times the next launch with the given tag, or nulls
// when profiling is off or tag is not a kernel tag (-1: not sampled).
void prof_events(Ctx &c, int tag, hipEvent_t *start, hipEvent_t *stop);
inline bool prof_sampled(const Ctx &c, int k) { return c.prof.on > 0 && (k % c.prof.on) == 0; }
// Folds finished records into the totals; the stream must be idle.
void prof_harvest(Ctx &c);

// Launch on the context's stream; a profiled launch goes through
// hipExtLaunchKernelGGL, which stamps the dispatch's own start and end times
// into the event pair (no marker packets around the kernel), so the figures
// are kernel durations, as a kernel trace reports them.
template <typename... KArgs, typename... Args>
inline void launch_tagged(Ctx &c, int tag, void (*kernel)(KArgs...), dim3 grid, dim3 block, Args... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  prof_events(c, tag, &e0, &e1);
  if (e0)
    hipExtLaunchKernelGGL(kernel, grid, block, 0, c.stream, e0, e1, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, 0, c.stream, args...);
}

inline __host__ __device__ long hidx(const Mesh &m, int i, int j) { return (long)j * (m.nx + 2) + i; }

// ------------------------------------------------------------------ launch helpers
#define COVOF_CHECK(expr)                                   \
  do {                                                      \
    hipError_t e_ = (expr);                                 \
    if (e_ != hipSuccess) return COVOF_HIP_ERROR;           \
  } while (0)

#define COVOF_CHECK_LAUNCH() COVOF_CHECK(hipGetLastError())

// Serial sums in the reference's element order (sum(x*y) / dot_product over
// column-major arrays: one accumulator, index order, each product rounded
// before it is added): sum q of a[q][t] * b[q][t] (b null: a[q][t] * z with
// z = Jacobi(a[q][t], dg[q][t]) when dg is set, else a[q][t] alone), written
// to dst[q][0], with dst[q][1 .. nzero-1] zeroed (a partial-sum slot whose
// grid total is then the serial sum exactly).  Nothing runs while the scalar
// slot f0 or f1 is nonzero (a stopped solve).
constexpr int SER_MAX = 5;
struct SerialSet {
  const double *a[SER_MAX], *b[SER_MAX], *dg[SER_MAX];
  double *dst[SER_MAX];
  int nq, nzero;
};
void serial_sums(Ctx &c, long n, const SerialSet &s, int f0, int f1);

int mg_check(const Ctx &c, int kind);
int mg_prepare(Ctx &c, bool fp32);
// Drops the captured solver cycles and the multigrid coarse operators (the
// next solve captures and coarsens again): after a solver option changes.
void solver_invalidate(Ctx &c);
// Work arrays made on a path's first use, made ahead (covof_ctx_reserve).
int mom_reserve(Ctx &c);
int mg_reserve(Ctx &c);
int sgs_reserve(Ctx &c);
int cg2_reserve(Ctx &c);
int surface_reserve(Ctx &c);
int vof_adjoint_reserve(Ctx &c);
// z = M^-1 r; f0, f1: the scalar slots of the solve's stop flags
// rz: when not null and the cycle's finest post pass is fused, that pass also
// forms the partials of r.z, folded into the nblocks entries at rz (the
// reduction grid's r.z slot); the return value says whether it did.
bool mg_apply64(Ctx &c, int kind, const double *r, double *z, int f0, int f1, double *rz = nullptr);
void mg_apply32(Ctx &c, int kind, const float *r, float *z, int f0, int f1);
// z = M^-1 r of an FP64 solve with the V-cycle in FP32 (r rounded, z widened)
bool mg_apply_fp32_cycle(Ctx &c, const double *r, double *z, int f0, int f1, double *rz = nullptr);
void mg_free(Ctx &c);

int vof_alloc(Ctx &c);
void vof_free(Ctx &c);
// surface.hip: curvature + CSF (G: the field the normals come from), the two
// smoothings of advect_vof; device pointers, halo layout, no aliasing of an
// output with an input
int curvature_csf(Ctx &c, const double *G, const double *C, int enable, double sigma, double *curv,
                  double *fx, double *fy, int method = COVOF_CURVATURE_NORMALS);
int smooth_diffusion(Ctx &c, const double *C, int steps, double *Ct);
int svof_smooth(Ctx &c, const double *C, double width, double *Cs);
void surface_free(Ctx &c);
int poisson_alloc(Ctx &c);
void poisson_free(Ctx &c);
int poisson_matvec(Ctx &c, const double *x, double *y);
// p(1:nx,1:ny) = x - mean(x) with the periodic halo wrap (end of the SP Poisson block)
int sp_pressure(Ctx &c, const double *x, double *p);
int poisson_solve(Ctx &c, const double *b, double *x, int lin_solver, int precond, double tol, int maxit,
                  int precision, int *iters, double *relres);
int poisson_precond(Ctx &c, int precond, const double *r, double *z);
// krylov.hip: BiCGSTAB(2) and CG of module_linear_solvers on any operator
int krylov_bicgstab2(Ctx &c, LinOp &op, const double *b, double *x, double *const w[9], double tol,
                     int maxit, int *iters, double *relres);
int krylov_cg(Ctx &c, LinOp &op, const double *b, double *x, double *const w[4], double tol, int maxit,
              int *iters, double *relres);
// momentum.hip: the SP momentum block
int mom_bind(Ctx &c, const MomCfg &cfg, const MomFields &f, double dt);
int mom_diag(Ctx &c, double *diag);
int mom_rhs(Ctx &c, double *rhs);
int mom_pack(Ctx &c, const double *u, const double *v, const double *p, double *x);
int mom_matvec(Ctx &c, const double *x, double *y);
int mom_solve(Ctx &c, const double *b, double *x, int lin_solver, int precond, double tol, int maxit,
              int *iters, double *relres);
int mom_predictor(Ctx &c, double *u, double *v, int lin_solver, int precond, double tol, int maxit,
                  int *iters, double *relres);
int mono_step(Ctx &c, double *u, double *v, double *p, int lin_solver, int precond, double tol, int maxit,
              int *iters, double *relres);
int mom_precond(Ctx &c, int precond, const double *r, double *z);
int mom_apply_precond(Ctx &c, int precond, const double *r, double *z);
void mom_free(Ctx &c);
int velocity_bc(Ctx &c, const MomCfg &cfg, double *u, double *v, int homogeneous);
// diagnostics.hip: output-side reductions
int history_row(Ctx &c, const double *C, const double *u, const double *v, const double *p, const double *rho,
                const double *T, double *out);
int cell_output(Ctx &c, const double *u, const double *v, double *div, double *uc, double *vc);

}  // namespace covof
