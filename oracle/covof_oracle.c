/*
 * CPU oracle of the COVOF hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference Fortran for the per-timestep hot path
 * (interface geometry, PLIC/ELVIRA reconstruction, Weymouth-Yue advection,
 * VOF halo/BC, physical properties, SP Poisson assembly / matvec / SGS / PCG,
 * BiCGSTAB(2), velocity correction).  Every function cites the reference
 * file:line it follows.  It is compiled with -ffp-contract=off and evaluates
 * every expression in the reference's left-to-right order, so that on
 * identical inputs it reproduces the reference bit-for-bit.
 *
 * Who may use it: tests/ (as the checker), __graft_entry__.smoke() (as the
 * checker) and bench.py's cpu_baseline leg (as the timed CPU baseline).  The
 * shipped library never links, loads or calls anything here.
 *
 * Pins (where each function is checked):
 *   cut_box_area, compute_alpha   golden vectors from the compiled reference
 *                                 (tests/test_oracle_golden.py); closed-form
 *                                 areas, quadrature of random cuts, bisection
 *                                 round trip and thresholds (test_oracle_pins.py)
 *   reconstruct_* / interface     golden vectors (PLIC, ELVIRA; periodic, walls,
 *                                 mixed); ELVIRA straight-line exactness and
 *                                 Youngs = gradient of a linear field (pins)
 *   wy_step, apply_vof_bc,        golden multi-step vectors (both Strang orders,
 *   update_properties             periodic and walls, with properties); volume
 *                                 telescoping and exact translation (pins)
 *   poisson_assemble / matvec     closed-form spectra (NEUMANN, periodic, walls),
 *                                 symmetry, SPD, the quadratic form of the face
 *                                 coefficients (pins)
 *   poisson_sgs / precond         the explicit (D-L)D^-1(D-U) solve on walls;
 *                                 Jacobi = r / a_P (pins)
 *   sp_pcg, bicgstab2             dense direct solves (pins)
 *   sp_finish_pressure,           discrete divergence of the corrected velocity
 *   sp_correct                    equals the solve residual (pins)
 *   bench_jacobi_pcg              timing loop only (the CPU baseline)
 */
#include "covof_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

enum { BC_DIRICHLET = 1, BC_NEUMANN = 2, BC_PERIODIC = 3, BC_FREESLIP = 4 };

#define H(m, i, j) ((long)(j) * ((m)->nx + 2) + (i))      /* halo field index  */
#define W(m, i, j) ((long)((j) - 1) * (m)->nx + ((i) - 1)) /* (nx,ny) work index */

/* src/module_mesh.f90:67-144 (single rank: i_start_global = 1, dx = lx/nx). */
void orc_mesh_init(orc_mesh *m, int nx, int ny, double lx, double ly, const int bc[4]) {
  m->nx = nx;
  m->ny = ny;
  m->lx = lx;
  m->ly = ly;
  m->dx = lx / (double)nx;
  m->dy = ly / (double)ny;
  for (int k = 0; k < 4; ++k) m->bc[k] = bc[k];
}

/* Periodicity of the Cartesian topology (src/module_simulation.f90:115-117):
 * on one rank a periodic dimension has no MPI_PROC_NULL neighbour, so the
 * is_boundary_* flags of src/module_mesh.f90:139-142 are the negation. */
static int per_x(const orc_mesh *m) { return m->bc[0] == BC_PERIODIC || m->bc[1] == BC_PERIODIC; }
static int per_y(const orc_mesh *m) { return m->bc[2] == BC_PERIODIC || m->bc[3] == BC_PERIODIC; }

/* src/module_mesh.f90:162-176 */
static double x_face(const orc_mesh *m, int i) { return (double)(i - 1) * m->dx; }
static double y_face(const orc_mesh *m, int j) { return (double)(j - 1) * m->dy; }

static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }

/* src/module_interface_reconstruction.f90:33-101 */
double orc_cut_box_area(double xmin, double xmax, double ymin, double ymax,
                        double nx, double ny, double alpha) {
  if ((xmax - xmin) <= 1.0e-14 || (ymax - ymin) <= 1.0e-14) return 0.0;
  const double bx[4] = {xmin, xmax, xmax, xmin};
  const double by[4] = {ymin, ymin, ymax, ymax};
  double d[4];
  for (int k = 0; k < 4; ++k) d[k] = nx * bx[k] + ny * by[k] + alpha;
  double cx[8], cy[8];
  int n = 0;
  for (int k = 0; k < 4; ++k) {
    const int kn = (k + 1) & 3;
    const double d1 = d[k], d2 = d[kn];
    if (d1 <= 0.0) {
      if (d2 <= 0.0) {
        cx[n] = bx[kn];
        cy[n] = by[kn];
        ++n;
      } else {
        const double t = -d1 / (d2 - d1);
        cx[n] = bx[k] + t * (bx[kn] - bx[k]);
        cy[n] = by[k] + t * (by[kn] - by[k]);
        ++n;
      }
    } else if (d2 <= 0.0) {
      const double t = -d1 / (d2 - d1);
      cx[n] = bx[k] + t * (bx[kn] - bx[k]);
      cy[n] = by[k] + t * (by[kn] - by[k]);
      ++n;
      cx[n] = bx[kn];
      cy[n] = by[kn];
      ++n;
    }
  }
  if (n < 3) return 0.0;
  double s = 0.0;
  for (int p = 0; p < n; ++p) {
    const int q = (p + 1) % n;
    s = s + (cx[p] * cy[q] - cx[q] * cy[p]);
  }
  return 0.5 * fabs(s);
}

/* src/module_interface_reconstruction.f90:105-147 */
double orc_compute_alpha(double xmin, double xmax, double ymin, double ymax,
                         double nx, double ny, double c) {
  const double total = (xmax - xmin) * (ymax - ymin);
  if (c <= 1.0e-7) return 1.0e9;
  if (c >= 1.0 - 1.0e-7) return -1.0e9;
  const double target = dmax(0.0, dmin(1.0, c)) * total;
  const double v1 = nx * xmin + ny * ymin;
  const double v2 = nx * xmax + ny * ymin;
  const double v3 = nx * xmax + ny * ymax;
  const double v4 = nx * xmin + ny * ymax;
  const double vmin = dmin(v1, dmin(v2, dmin(v3, v4)));
  const double vmax = dmax(v1, dmax(v2, dmax(v3, v4)));
  double lo = -vmax - 1.0e-12;
  double hi = -vmin + 1.0e-12;
  for (int it = 0; it < 45; ++it) {
    const double mid = 0.5 * (lo + hi);
    const double a = orc_cut_box_area(xmin, xmax, ymin, ymax, nx, ny, mid);
    if (a < target)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

/* Threads of the cell loops of the reconstruction and the WY pass (1: serial,
 * as the tests run them; the loops' iterations are independent, so the
 * result is the same for any count).  bench.py's CPU baseline raises it. */
static int g_threads = 1;
void orc_set_threads(int t) { g_threads = t > 1 ? t : 1; }

static int interior(const orc_mesh *m, int i, int j) {
  return i >= 1 && i <= m->nx && j >= 1 && j <= m->ny;
}

/* src/module_interface_reconstruction.f90:283-343 (Youngs) */
static void plic_cell(const orc_mesh *m, const double *C, int i, int j,
                      double *onx, double *ony, double *oal) {
  const double c = C[H(m, i, j)];
  if (c <= 1.0e-6) { *onx = 1.0; *ony = 0.0; *oal = 1.0e9; return; }
  if (c >= 1.0 - 1.0e-6) { *onx = 1.0; *ony = 0.0; *oal = -1.0e9; return; }
  double mx, my;
  if (interior(m, i, j)) {
    const double c_mm = C[H(m, i - 1, j - 1)], c_0m = C[H(m, i, j - 1)], c_pm = C[H(m, i + 1, j - 1)];
    const double c_m0 = C[H(m, i - 1, j)], c_p0 = C[H(m, i + 1, j)];
    const double c_mp = C[H(m, i - 1, j + 1)], c_0p = C[H(m, i, j + 1)], c_pp = C[H(m, i + 1, j + 1)];
    mx = -(((c_pp + 2.0 * c_p0 + c_pm) - (c_mp + 2.0 * c_m0 + c_mm)) / (8.0 * m->dx));
    my = -(((c_pp + 2.0 * c_0p + c_mp) - (c_pm + 2.0 * c_0m + c_mm)) / (8.0 * m->dy));
  } else {
    mx = 1.0;
    my = 0.0;
  }
  const double nrm = hypot(mx, my);
  double pnx, pny;
  if (nrm > 1.0e-12) { pnx = mx / nrm; pny = my / nrm; } else { pnx = 1.0; pny = 0.0; }
  *onx = pnx;
  *ony = pny;
  *oal = orc_compute_alpha(x_face(m, i), x_face(m, i + 1), y_face(m, j), y_face(m, j + 1), pnx, pny, c);
}

/* src/module_interface_reconstruction.f90:346-457 (ELVIRA, 7 candidates) */
static void elvira_cell(const orc_mesh *m, const double *C, int i, int j,
                        double *onx, double *ony, double *oal) {
  const double c = C[H(m, i, j)];
  if (c <= 1.0e-6) { *onx = 1.0; *ony = 0.0; *oal = 1.0e9; return; }
  if (c >= 1.0 - 1.0e-6) { *onx = 1.0; *ony = 0.0; *oal = -1.0e9; return; }
  const double dx = m->dx, dy = m->dy;
  const double xmin = x_face(m, i), xmax = x_face(m, i + 1);
  const double ymin = y_face(m, j), ymax = y_face(m, j + 1);
  if (!interior(m, i, j)) {
    *onx = 1.0;
    *ony = 0.0;
    *oal = orc_compute_alpha(xmin, xmax, ymin, ymax, 1.0, 0.0, c);
    return;
  }
#define CC(a, b) C[H(m, i + (a), j + (b))]
  const double H_m1 = (CC(-1, -1) + CC(-1, 0) + CC(-1, 1)) * dy;
  const double H_0 = (CC(0, -1) + CC(0, 0) + CC(0, 1)) * dy;
  const double H_p1 = (CC(1, -1) + CC(1, 0) + CC(1, 1)) * dy;
  const double W_m1 = (CC(-1, -1) + CC(0, -1) + CC(1, -1)) * dx;
  const double W_0 = (CC(-1, 0) + CC(0, 0) + CC(1, 0)) * dx;
  const double W_p1 = (CC(-1, 1) + CC(0, 1) + CC(1, 1)) * dx;
  const double sgn_y = (CC(0, 1) < CC(0, -1)) ? 1.0 : -1.0;
  const double sgn_x = (CC(1, 0) < CC(-1, 0)) ? 1.0 : -1.0;
  double cnx[7], cny[7];
  cnx[0] = -(sgn_y * (H_p1 - H_m1) / (2.0 * dx)); cny[0] = sgn_y * 1.0;
  cnx[1] = -(sgn_y * (H_p1 - H_0) / dx);          cny[1] = sgn_y * 1.0;
  cnx[2] = -(sgn_y * (H_0 - H_m1) / dx);          cny[2] = sgn_y * 1.0;
  cnx[3] = sgn_x * 1.0; cny[3] = -(sgn_x * (W_p1 - W_m1) / (2.0 * dy));
  cnx[4] = sgn_x * 1.0; cny[4] = -(sgn_x * (W_p1 - W_0) / dy);
  cnx[5] = sgn_x * 1.0; cny[5] = -(sgn_x * (W_0 - W_m1) / dy);
  cnx[6] = -(((CC(1, 1) + 2.0 * CC(1, 0) + CC(1, -1)) - (CC(-1, 1) + 2.0 * CC(-1, 0) + CC(-1, -1))) / (8.0 * dx));
  cny[6] = -(((CC(1, 1) + 2.0 * CC(0, 1) + CC(-1, 1)) - (CC(1, -1) + 2.0 * CC(0, -1) + CC(-1, -1))) / (8.0 * dy));
  double best_err = 1.0e30, bnx = 1.0, bny = 0.0, bal = 0.0;
  for (int k = 0; k < 7; ++k) {
    const double cn = hypot(cnx[k], cny[k]);
    if (cn < 1.0e-12) continue;
    const double ux = cnx[k] / cn, uy = cny[k] / cn;
    const double al = orc_compute_alpha(xmin, xmax, ymin, ymax, ux, uy, c);
    double err = 0.0;
    for (int dj = -1; dj <= 1; ++dj)
      for (int di = -1; di <= 1; ++di) {
        if (di == 0 && dj == 0) continue;
        const double vt = CC(di, dj) * dx * dy;
        const double vr = orc_cut_box_area(x_face(m, i + di), x_face(m, i + di + 1),
                                           y_face(m, j + dj), y_face(m, j + dj + 1), ux, uy, al);
        const double df = vr - vt;
        err = err + df * df;
      }
    if (err < best_err) { best_err = err; bnx = ux; bny = uy; bal = al; }
  }
#undef CC
  *onx = bnx;
  *ony = bny;
  *oal = bal;
}

void orc_reconstruct(const orc_mesh *m, int method, const double *C,
                     double *snx, double *sny, double *sal) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads) if (g_threads > 1)
  for (int j = 0; j <= m->ny + 1; ++j)
    for (int i = 0; i <= m->nx + 1; ++i) {
      const long k = H(m, i, j);
      if (method == 1)
        elvira_cell(m, C, i, j, &snx[k], &sny[k], &sal[k]);
      else
        plic_cell(m, C, i, j, &snx[k], &sny[k], &sal[k]);
    }
}

/* src/module_mpi_comm.f90:46-56 (single-rank branch) */
void orc_exchange_halos_scalar(const orc_mesh *m, double *f) {
  const int nx = m->nx, ny = m->ny;
  if (per_x(m)) {
    for (int j = 0; j <= ny + 1; ++j) f[H(m, 0, j)] = f[H(m, nx, j)];
    for (int j = 0; j <= ny + 1; ++j) f[H(m, nx + 1, j)] = f[H(m, 1, j)];
  }
  if (per_y(m)) {
    for (int i = 0; i <= nx + 1; ++i) f[H(m, i, 0)] = f[H(m, i, ny)];
    for (int i = 0; i <= nx + 1; ++i) f[H(m, i, ny + 1)] = f[H(m, i, 1)];
  }
}

/* src/module_interface_reconstruction.f90:580-683 */
void orc_reconstruct_interface(const orc_mesh *m, int method, const double *C,
                               double *snx, double *sny, double *sal) {
  const int nx = m->nx, ny = m->ny;
  const long n = (long)(nx + 2) * (ny + 2);
  orc_reconstruct(m, method, C, snx, sny, sal);
  double *gx = calloc(n, sizeof(double)), *gy = calloc(n, sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      gx[H(m, i, j)] = snx[H(m, i, j)];
      gy[H(m, i, j)] = sny[H(m, i, j)];
    }
  orc_exchange_halos_scalar(m, gx);
  orc_exchange_halos_scalar(m, gy);
  if (!per_x(m)) {
    for (int j = 1; j <= ny; ++j) { gx[H(m, 0, j)] = gx[H(m, 1, j)]; gy[H(m, 0, j)] = gy[H(m, 1, j)]; }
    for (int j = 1; j <= ny; ++j) { gx[H(m, nx + 1, j)] = gx[H(m, nx, j)]; gy[H(m, nx + 1, j)] = gy[H(m, nx, j)]; }
  }
  if (!per_y(m)) {
    for (int i = 0; i <= nx + 1; ++i) { gx[H(m, i, 0)] = gx[H(m, i, 1)]; gy[H(m, i, 0)] = gy[H(m, i, 1)]; }
    for (int i = 0; i <= nx + 1; ++i) { gx[H(m, i, ny + 1)] = gx[H(m, i, ny)]; gy[H(m, i, ny + 1)] = gy[H(m, i, ny)]; }
  }
#define FIX(ii, jj)                                                                  \
  do {                                                                               \
    const long k_ = H(m, ii, jj);                                                    \
    const double c_ = C[k_];                                                         \
    if (c_ > 1.0e-6 && c_ < 1.0 - 1.0e-6) {                                          \
      double sx_ = gx[k_], sy_ = gy[k_];                                             \
      const double nr_ = hypot(sx_, sy_);                                            \
      if (nr_ > 1.0e-12) {                                                           \
        sx_ = sx_ / nr_;                                                             \
        sy_ = sy_ / nr_;                                                             \
        snx[k_] = sx_;                                                               \
        sny[k_] = sy_;                                                               \
        sal[k_] = orc_compute_alpha(x_face(m, ii), x_face(m, (ii) + 1), y_face(m, jj), \
                                    y_face(m, (jj) + 1), sx_, sy_, c_);              \
      }                                                                              \
    }                                                                                \
  } while (0)
  for (int j = 1; j <= ny; ++j) {
    FIX(0, j);
    FIX(nx + 1, j);
  }
  for (int i = 0; i <= nx + 1; ++i) {
    FIX(i, 0);
    FIX(i, ny + 1);
  }
#undef FIX
  free(gx);
  free(gy);
}

/* The top side's JET_INLET profile (cfg%bc_top%profile, jet_center,
 * jet_width); off unless orc_set_top_jet turns it on. */
static struct {
  int on;
  double center, width;
} g_jet = {0, 0.5, 0.08};

void orc_set_top_jet(int on, double center, double width) {
  g_jet.on = on;
  g_jet.center = center;
  g_jet.width = width;
}

/* src/module_mpi_comm.f90:636-683 (single rank) */
void orc_apply_vof_bc(const orc_mesh *m, double *C) {
  const int nx = m->nx, ny = m->ny;
  if (m->bc[0] == BC_PERIODIC) {
    for (int j = 1; j <= ny; ++j) C[H(m, 0, j)] = C[H(m, nx, j)];
    for (int j = 1; j <= ny; ++j) C[H(m, nx + 1, j)] = C[H(m, 1, j)];
  } else {
    if (!per_x(m)) for (int j = 1; j <= ny; ++j) C[H(m, 0, j)] = C[H(m, 1, j)];
    if (!per_x(m)) for (int j = 1; j <= ny; ++j) C[H(m, nx + 1, j)] = C[H(m, nx, j)];
  }
  if (m->bc[2] == BC_PERIODIC) {
    for (int i = 1; i <= nx; ++i) C[H(m, i, 0)] = C[H(m, i, ny)];
    for (int i = 1; i <= nx; ++i) C[H(m, i, ny + 1)] = C[H(m, i, 1)];
  } else {
    if (!per_y(m)) for (int i = 1; i <= nx; ++i) C[H(m, i, 0)] = C[H(m, i, 1)];
  }
  if (!per_y(m)) {
    if (g_jet.on) {
      const double xc_top = g_jet.center > 0.0 ? g_jet.center : 0.5 * m->lx;
      const double half_w = g_jet.width > 0.0 ? 0.5 * g_jet.width : 0.04;
      for (int i = 1; i <= nx; ++i) {
        const double x_c = ((double)(i - 1) + 0.5) * m->dx;
        C[H(m, i, ny + 1)] = (fabs(x_c - xc_top) <= half_w) ? 0.0 : 1.0;
      }
    } else {
      for (int i = 1; i <= nx; ++i) C[H(m, i, ny + 1)] = C[H(m, i, ny)];
    }
  }
}

/* One directional pass of src/module_interface_vof.f90:1162-1377.
 * dir 0 = X (Eq. 15), 1 = Y (Eq. 16).  src -> dst (dst's halo starts as src's). */
static void wy_pass(const orc_mesh *m, int method, int dir, int clamp, const double *src,
                    double *dst, const double *u, const double *v, const double *cd,
                    double dt, double *flux, double *snx, double *sny, double *sal) {
  const int nx = m->nx, ny = m->ny;
  const long n = (long)(nx + 2) * (ny + 2);
  const double dx = m->dx, dy = m->dy, area = dx * dy;
  orc_reconstruct_interface(m, method, src, snx, sny, sal);
  memset(flux, 0, n * sizeof(double));
  if (dir == 0) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads) if (g_threads > 1)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx + 1; ++i) {
        const double uf = u[H(m, i, j)];
        const double del = fabs(uf) * dt;
        if (del <= 1.0e-14) continue;
        const double xp = x_face(m, i);
        if (uf > 0.0) {
          const long k = H(m, i - 1, j);
          flux[H(m, i, j)] = orc_cut_box_area(xp - del, xp, y_face(m, j), y_face(m, j + 1), snx[k], sny[k], sal[k]);
        } else {
          const long k = H(m, i, j);
          flux[H(m, i, j)] = -orc_cut_box_area(xp, xp + del, y_face(m, j), y_face(m, j + 1), snx[k], sny[k], sal[k]);
        }
      }
    if (m->bc[0] == BC_PERIODIC)
      for (int j = 1; j <= ny; ++j) {
        if (u[H(m, 1, j)] >= 0.0)
          flux[H(m, 1, j)] = flux[H(m, nx + 1, j)];
        else
          flux[H(m, nx + 1, j)] = flux[H(m, 1, j)];
      }
    memcpy(dst, src, n * sizeof(double));
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        double c = src[H(m, i, j)] - (flux[H(m, i + 1, j)] - flux[H(m, i, j)]) / area +
                   cd[H(m, i, j)] * (dt / dx) * (u[H(m, i + 1, j)] - u[H(m, i, j)]);
        if (clamp) c = dmax(0.0, dmin(1.0, c));
        dst[H(m, i, j)] = c;
      }
  } else {
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads) if (g_threads > 1)
    for (int j = 1; j <= ny + 1; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double vf = v[H(m, i, j)];
        const double del = fabs(vf) * dt;
        if (del <= 1.0e-14) continue;
        const double yp = y_face(m, j);
        if (vf > 0.0) {
          const long k = H(m, i, j - 1);
          flux[H(m, i, j)] = orc_cut_box_area(x_face(m, i), x_face(m, i + 1), yp - del, yp, snx[k], sny[k], sal[k]);
        } else {
          const long k = H(m, i, j);
          flux[H(m, i, j)] = -orc_cut_box_area(x_face(m, i), x_face(m, i + 1), yp, yp + del, snx[k], sny[k], sal[k]);
        }
      }
    if (m->bc[2] == BC_PERIODIC)
      for (int i = 1; i <= nx; ++i) {
        if (v[H(m, i, 1)] >= 0.0)
          flux[H(m, i, 1)] = flux[H(m, i, ny + 1)];
        else
          flux[H(m, i, ny + 1)] = flux[H(m, i, 1)];
      }
    memcpy(dst, src, n * sizeof(double));
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        double c = src[H(m, i, j)] - (flux[H(m, i, j + 1)] - flux[H(m, i, j)]) / area +
                   cd[H(m, i, j)] * (dt / dy) * (v[H(m, i, j + 1)] - v[H(m, i, j)]);
        if (clamp) c = dmax(0.0, dmin(1.0, c));
        dst[H(m, i, j)] = c;
      }
  }
}

/* advect_vof with WEYMOUTH_YUE (src/module_interface_vof.f90:792-825, :1162-1377),
 * without the property update (orc_update_properties is the next call there). */
void orc_wy_step(const orc_mesh *m, int method, double *C, const double *u, const double *v,
                 double dt, int *step_count) {
  const long n = (long)(m->nx + 2) * (m->ny + 2);
  double *cd = malloc(n * sizeof(double)), *cs = malloc(n * sizeof(double));
  double *fl = malloc(n * sizeof(double)), *sx = malloc(n * sizeof(double));
  double *sy = malloc(n * sizeof(double)), *sa = malloc(n * sizeof(double));
  for (long k = 0; k < n; ++k) cd[k] = (C[k] >= 0.5) ? 1.0 : 0.0;
  const int x_first = ((*step_count) % 2) == 0;
  const int d1 = x_first ? 0 : 1, d2 = x_first ? 1 : 0;
  wy_pass(m, method, d1, 0, C, cs, u, v, cd, dt, fl, sx, sy, sa);
  orc_exchange_halos_scalar(m, cs);
  orc_apply_vof_bc(m, cs);
  wy_pass(m, method, d2, 1, cs, C, u, v, cd, dt, fl, sx, sy, sa);
  *step_count += 1;
  orc_exchange_halos_scalar(m, C);
  orc_apply_vof_bc(m, C);
  free(cd); free(cs); free(fl); free(sx); free(sy); free(sa);
}

/* src/module_interface_vof.f90:479-616 (incompressible); Cs non-null: the
 * SVOF branch (rho, mu from C_s, the harmonic node viscosity from the node
 * mean of C_s, :540-545, :596-600). */
void orc_update_properties_cs(const orc_mesh *m, double *C, const double *Cs, double rho1, double rho2,
                              double mu1, double mu2, double *rho, double *rho_u, double *rho_v, double *mu,
                              double *mu_node) {
  const int nx = m->nx, ny = m->ny;
  const int bl = !per_x(m), br = !per_x(m), bb = !per_y(m), bt = !per_y(m);
  if (bl && bb) C[H(m, 0, 0)] = C[H(m, 1, 1)];
  if (br && bb) C[H(m, nx + 1, 0)] = C[H(m, nx, 1)];
  if (bl && bt) C[H(m, 0, ny + 1)] = C[H(m, 1, ny)];
  if (br && bt) C[H(m, nx + 1, ny + 1)] = C[H(m, nx, ny)];
  const double *F = Cs ? Cs : C;
  for (int j = 0; j <= ny + 1; ++j)
    for (int i = 0; i <= nx + 1; ++i) {
      const double c = dmax(0.0, dmin(1.0, F[H(m, i, j)]));
      rho[H(m, i, j)] = c * rho1 + (1.0 - c) * rho2;
      mu[H(m, i, j)] = c * mu1 + (1.0 - c) * mu2;
    }
  mu[H(m, 0, 0)] = mu[H(m, 1, 1)];
  mu[H(m, nx + 1, 0)] = mu[H(m, nx, 1)];
  mu[H(m, 0, ny + 1)] = mu[H(m, 1, ny)];
  mu[H(m, nx + 1, ny + 1)] = mu[H(m, nx, ny)];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx + 1; ++i) rho_u[H(m, i, j)] = 0.5 * (rho[H(m, i - 1, j)] + rho[H(m, i, j)]);
  for (int j = 1; j <= ny + 1; ++j)
    for (int i = 1; i <= nx; ++i) rho_v[H(m, i, j)] = 0.5 * (rho[H(m, i, j - 1)] + rho[H(m, i, j)]);
  for (int j = 1; j <= ny + 1; ++j)
    for (int i = 1; i <= nx + 1; ++i) {
      if (Cs) {
        double cn = 0.25 * (Cs[H(m, i - 1, j - 1)] + Cs[H(m, i, j - 1)] + Cs[H(m, i - 1, j)] + Cs[H(m, i, j)]);
        cn = dmax(0.0, dmin(1.0, cn));
        const double den = mu2 * cn + mu1 * (1.0 - cn);
        mu_node[H(m, i, j)] = (mu1 * mu2) / dmax(1.0e-30, den);
      } else {
        const double sw = mu[H(m, i - 1, j - 1)], se = mu[H(m, i, j - 1)];
        const double nw = mu[H(m, i - 1, j)], ne = mu[H(m, i, j)];
        mu_node[H(m, i, j)] = 4.0 / (1.0 / sw + 1.0 / se + 1.0 / nw + 1.0 / ne);
      }
    }
}

/* src/module_interface_vof.f90:479-616 (incompressible, no SVOF smoothing). */
void orc_update_properties(const orc_mesh *m, double *C, double rho1, double rho2,
                           double mu1, double mu2, double *rho, double *rho_u,
                           double *rho_v, double *mu, double *mu_node) {
  const int nx = m->nx, ny = m->ny;
  const int bl = !per_x(m), br = !per_x(m), bb = !per_y(m), bt = !per_y(m);
  if (bl && bb) C[H(m, 0, 0)] = C[H(m, 1, 1)];
  if (br && bb) C[H(m, nx + 1, 0)] = C[H(m, nx, 1)];
  if (bl && bt) C[H(m, 0, ny + 1)] = C[H(m, 1, ny)];
  if (br && bt) C[H(m, nx + 1, ny + 1)] = C[H(m, nx, ny)];
  for (int j = 0; j <= ny + 1; ++j)
    for (int i = 0; i <= nx + 1; ++i) {
      const double c = dmax(0.0, dmin(1.0, C[H(m, i, j)]));
      rho[H(m, i, j)] = c * rho1 + (1.0 - c) * rho2;
      mu[H(m, i, j)] = c * mu1 + (1.0 - c) * mu2;
    }
  mu[H(m, 0, 0)] = mu[H(m, 1, 1)];
  mu[H(m, nx + 1, 0)] = mu[H(m, nx, 1)];
  mu[H(m, 0, ny + 1)] = mu[H(m, 1, ny)];
  mu[H(m, nx + 1, ny + 1)] = mu[H(m, nx, ny)];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx + 1; ++i) rho_u[H(m, i, j)] = 0.5 * (rho[H(m, i - 1, j)] + rho[H(m, i, j)]);
  for (int j = 1; j <= ny + 1; ++j)
    for (int i = 1; i <= nx; ++i) rho_v[H(m, i, j)] = 0.5 * (rho[H(m, i, j - 1)] + rho[H(m, i, j)]);
  for (int j = 1; j <= ny + 1; ++j)
    for (int i = 1; i <= nx + 1; ++i) {
      const double sw = mu[H(m, i - 1, j - 1)], se = mu[H(m, i, j - 1)];
      const double nw = mu[H(m, i - 1, j)], ne = mu[H(m, i, j)];
      mu_node[H(m, i, j)] = 4.0 / (1.0 / sw + 1.0 / se + 1.0 / nw + 1.0 / ne);
    }
}

/* ------------------------------------------------- smoothing, curvature, CSF */

static double clamp01(double c) {
  const double t = 1.0 < c ? 1.0 : c;
  return 0.0 > t ? 0.0 : t;
}

/* Homogeneous Neumann copies on the physical sides (interior rows / columns). */
static void neumann_sides(const orc_mesh *m, double *f) {
  const int nx = m->nx, ny = m->ny;
  for (int j = 1; j <= ny; ++j) {
    if (!per_x(m)) f[H(m, 0, j)] = f[H(m, 1, j)];
    if (!per_x(m)) f[H(m, nx + 1, j)] = f[H(m, nx, j)];
  }
  for (int i = 1; i <= nx; ++i) {
    if (!per_y(m)) f[H(m, i, 0)] = f[H(m, i, 1)];
    if (!per_y(m)) f[H(m, i, ny + 1)] = f[H(m, i, ny)];
  }
}

/* One lexicographic Gauss-Seidel sweep of (I - D Lap) Cs = C (forward or
 * backward), zero weights across physical sides (:413-449). */
static void svof_sweep(const orc_mesh *m, const double *C, double *Cs, double wx, double wy, int backward) {
  const int nx = m->nx, ny = m->ny;
  const int bl = !per_x(m), br = !per_x(m), bb = !per_y(m), bt = !per_y(m);
  for (int jj = 1; jj <= ny; ++jj)
    for (int ii = 1; ii <= nx; ++ii) {
      const int i = backward ? nx + 1 - ii : ii, j = backward ? ny + 1 - jj : jj;
      const double aE = (i == nx && br) ? 0.0 : wx;
      const double aW = (i == 1 && bl) ? 0.0 : wx;
      const double aN = (j == ny && bt) ? 0.0 : wy;
      const double aS = (j == 1 && bb) ? 0.0 : wy;
      const double aP = 1.0 + aE + aW + aN + aS;
      const double sum = aE * Cs[H(m, i + 1, j)] + aW * Cs[H(m, i - 1, j)] + aN * Cs[H(m, i, j + 1)] +
                         aS * Cs[H(m, i, j - 1)];
      Cs[H(m, i, j)] = (C[H(m, i, j)] + sum) / aP;
    }
}

/* src/module_interface_vof.f90:387-476 */
void orc_svof_smooth(const orc_mesh *m, const double *C, double width, double *Cs) {
  const int nx = m->nx, ny = m->ny;
  const long n = (long)(nx + 2) * (ny + 2);
  double k = width;
  if (k <= 0.0) k = 1.0;
  const double D = k * m->dx * m->dy;
  const double wx = D / (m->dx * m->dx), wy = D / (m->dy * m->dy);
  memcpy(Cs, C, n * sizeof(double));
  for (int it = 0; it < 25; ++it) {
    neumann_sides(m, Cs);
    orc_exchange_halos_scalar(m, Cs);
    svof_sweep(m, C, Cs, wx, wy, 0);
    svof_sweep(m, C, Cs, wx, wy, 1);
  }
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) Cs[H(m, i, j)] = clamp01(Cs[H(m, i, j)]);
  neumann_sides(m, Cs);
  Cs[H(m, 0, 0)] = Cs[H(m, 1, 1)];
  Cs[H(m, nx + 1, 0)] = Cs[H(m, nx, 1)];
  Cs[H(m, 0, ny + 1)] = Cs[H(m, 1, ny)];
  Cs[H(m, nx + 1, ny + 1)] = Cs[H(m, nx, ny)];
  orc_exchange_halos_scalar(m, Cs);
}

/* Halo of vof_smooth_diffusion: exchange, then zero-gradient copies on the
 * physical sides (rows 0 / ny+1 over the full width, :1436-1440). */
static void diffusion_halo(const orc_mesh *m, double *f) {
  const int nx = m->nx, ny = m->ny;
  orc_exchange_halos_scalar(m, f);
  if (!per_x(m)) {
    for (int j = 1; j <= ny; ++j) f[H(m, 0, j)] = f[H(m, 1, j)];
    for (int j = 1; j <= ny; ++j) f[H(m, nx + 1, j)] = f[H(m, nx, j)];
  }
  if (!per_y(m)) {
    for (int i = 0; i <= nx + 1; ++i) f[H(m, i, 0)] = f[H(m, i, 1)];
    for (int i = 0; i <= nx + 1; ++i) f[H(m, i, ny + 1)] = f[H(m, i, ny)];
  }
}

/* src/module_interface_vof.f90:1416-1461 */
void orc_smooth_diffusion(const orc_mesh *m, const double *C, int steps, double *Ct) {
  const int nx = m->nx, ny = m->ny;
  const long n = (long)(nx + 2) * (ny + 2);
  const double dx = m->dx, dy = m->dy;
  memcpy(Ct, C, n * sizeof(double));
  const double dtau = 0.2 * dmin(dx * dx, dy * dy);
  if (steps < 1) steps = 3;
  double *cn = malloc(n * sizeof(double));
  for (int s = 0; s < steps; ++s) {
    diffusion_halo(m, Ct);
    memcpy(cn, Ct, n * sizeof(double));
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double c = Ct[H(m, i, j)];
        const double lx = (Ct[H(m, i + 1, j)] - 2.0 * c + Ct[H(m, i - 1, j)]) / (dx * dx);
        const double ly = (Ct[H(m, i, j + 1)] - 2.0 * c + Ct[H(m, i, j - 1)]) / (dy * dy);
        cn[H(m, i, j)] = c + dtau * (lx + ly);
      }
    memcpy(Ct, cn, n * sizeof(double));
  }
  diffusion_halo(m, Ct);
  free(cn);
}

/* ---- height-function curvature (no reference counterpart: the method of
 * Cummins, Francois & Kothe, Computers & Structures 83 (2005) 425-434, 3x7
 * column heights from the volume fractions).  C at (ii, jj) for the stencils:
 * the periodic image, else the nearest interior cell. */
static double hf_c(const orc_mesh *m, const double *C, int ii, int jj) {
  const int nx = m->nx, ny = m->ny;
  if (ii < 1) ii = per_x(m) ? ii + nx : 1;
  else if (ii > nx) ii = per_x(m) ? ii - nx : nx;
  if (jj < 1) jj = per_y(m) ? jj + ny : 1;
  else if (jj > ny) jj = per_y(m) ? jj - ny : ny;
  return C[H(m, ii, jj)];
}

/* Curvature at (i, j) from the heights along the dominant normal direction
 * (nv: the unit normal, towards increasing C): three 7-cell column sums,
 * kappa = -h'' / (1 + h'^2)^(3/2) (with the reference's sign: +1/R for a
 * liquid drop).  Returns 0 when a column does not run from empty (C < 1e-6)
 * to full (C > 1 - 1e-6) along the normal. */
int orc_hf_kappa(const orc_mesh *m, const double *C, int i, int j, double nxv, double nyv, double *kap) {
  const int vert = fabs(nyv) >= fabs(nxv);
  const double sgn = vert ? nyv : nxv;
  double h[3];
  for (int q = -1; q <= 1; ++q) {
    double sum = 0.0;
    for (int r = -3; r <= 3; ++r) sum = sum + (vert ? hf_c(m, C, i + q, j + r) : hf_c(m, C, i + r, j + q));
    const double lo = vert ? hf_c(m, C, i + q, j - 3) : hf_c(m, C, i - 3, j + q);
    const double hi = vert ? hf_c(m, C, i + q, j + 3) : hf_c(m, C, i + 3, j + q);
    if (sgn > 0.0 ? (lo > 1.0e-6 || hi < 1.0 - 1.0e-6) : (lo < 1.0 - 1.0e-6 || hi > 1.0e-6)) return 0;
    h[q + 1] = sum;
  }
  const double a = vert ? m->dy : m->dx, b = vert ? m->dx : m->dy;  /* along the heights / across */
  const double hp = (h[2] - h[0]) * a / (2.0 * b);
  const double hpp = (h[2] - 2.0 * h[1] + h[0]) * a / (b * b);
  const double den = 1.0 + hp * hp;
  *kap = -hpp / (den * sqrt(den));
  return 1;
}

/* src/module_interface_vof.f90:627-789 (no markers); method 1: the curvature
 * from heights of C where they bracket the interface (orc_hf_kappa), the
 * reference's -div(n) elsewhere. */
void orc_curvature_csf2(const orc_mesh *m, const double *G, const double *C, int enable, double sigma,
                        int method, double *curv, double *fx, double *fy) {
  const int nx = m->nx, ny = m->ny;
  const long n = (long)(nx + 2) * (ny + 2);
  const double dx = m->dx, dy = m->dy;
  if (!enable || sigma <= 0.0) {
    for (long k = 0; k < n; ++k) curv[k] = fx[k] = fy[k] = 0.0;
    return;
  }
  double *nxs = calloc(n, sizeof(double)), *nys = calloc(n, sizeof(double)), *kap = calloc(n, sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double gx = (G[H(m, i + 1, j)] - G[H(m, i - 1, j)]) / (2.0 * dx);
      const double gy = (G[H(m, i, j + 1)] - G[H(m, i, j - 1)]) / (2.0 * dy);
      const double gm = sqrt(gx * gx + gy * gy);
      if (gm > 1.0e-3 / dx) {
        nxs[H(m, i, j)] = gx / gm;
        nys[H(m, i, j)] = gy / gm;
      }
    }
  orc_exchange_halos_scalar(m, nxs);
  orc_exchange_halos_scalar(m, nys);
  unsigned char *ok = calloc(n, 1);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      if (nxs[H(m, i, j)] != 0.0 || nys[H(m, i, j)] != 0.0) {
        double k = -((nxs[H(m, i + 1, j)] - nxs[H(m, i - 1, j)]) / (2.0 * dx) +
                     (nys[H(m, i, j + 1)] - nys[H(m, i, j - 1)]) / (2.0 * dy));
        double khf;
        if (method == 1 && orc_hf_kappa(m, C, i, j, nxs[H(m, i, j)], nys[H(m, i, j)], &khf)) {
          k = khf;
          ok[H(m, i, j)] = 1;
        }
        kap[H(m, i, j)] = k;
      }
  if (method == 1) {
    /* a cell whose heights do not bracket the interface takes the mean of
     * its valid neighbours' height-function curvatures (rows j-1..j+1, i
     * fastest; periodic images, no cells beyond a physical side), else keeps
     * the -div(n) value */
    double *kf = malloc(n * sizeof(double));
    memcpy(kf, kap, n * sizeof(double));
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        if (ok[H(m, i, j)] || (nxs[H(m, i, j)] == 0.0 && nys[H(m, i, j)] == 0.0)) continue;
        double sum = 0.0;
        int cnt = 0;
        for (int b = -1; b <= 1; ++b)
          for (int a = -1; a <= 1; ++a) {
            int ii = i + a, jj = j + b;
            if (ii < 1 || ii > nx) {
              if (!per_x(m)) continue;
              ii = ii < 1 ? ii + nx : ii - nx;
            }
            if (jj < 1 || jj > ny) {
              if (!per_y(m)) continue;
              jj = jj < 1 ? jj + ny : jj - ny;
            }
            if (ok[H(m, ii, jj)]) {
              sum = sum + kap[H(m, ii, jj)];
              ++cnt;
            }
          }
        if (cnt > 0) kf[H(m, i, j)] = sum / (double)cnt;
      }
    memcpy(kap, kf, n * sizeof(double));
    free(kf);
  }
  free(ok);
  orc_exchange_halos_scalar(m, kap);
  memcpy(curv, kap, n * sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx + 1; ++i) {
      const double ka = 0.5 * (kap[H(m, i - 1, j)] + kap[H(m, i, j)]);
      const double gc = (C[H(m, i, j)] - C[H(m, i - 1, j)]) / dx;
      fx[H(m, i, j)] = sigma * ka * gc;
    }
  for (int j = 1; j <= ny + 1; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double ka = 0.5 * (kap[H(m, i, j - 1)] + kap[H(m, i, j)]);
      const double gc = (C[H(m, i, j)] - C[H(m, i, j - 1)]) / dy;
      fy[H(m, i, j)] = sigma * ka * gc;
    }
  free(nxs);
  free(nys);
  free(kap);
}

void orc_curvature_csf(const orc_mesh *m, const double *G, const double *C, int enable, double sigma,
                       double *curv, double *fx, double *fy) {
  orc_curvature_csf2(m, G, C, enable, sigma, 0, curv, fx, fy);
}

/* ---------------------------------------------------------------- Poisson */

static int wall(int btype) { return btype == BC_DIRICHLET || btype == BC_FREESLIP; }

/* The solid obstacle of config_t (obs_enabled, obs_type, box corners, disc
 * centre and radius, obs_perm); off unless orc_set_obstacle turns it on. */
static struct {
  int enabled, type;
  double x0, y0, x1, y1, xc, yc, r, perm;
} g_obs = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1.0e-40};

void orc_set_obstacle(int enabled, int type, double x0, double y0, double x1, double y1, double xc, double yc,
                      double r, double perm) {
  g_obs.enabled = enabled;
  g_obs.type = type;
  g_obs.x0 = x0; g_obs.y0 = y0; g_obs.x1 = x1; g_obs.y1 = y1;
  g_obs.xc = xc; g_obs.yc = yc; g_obs.r = r; g_obs.perm = perm;
}

/* is_in_obstacle (src/module_mesh.f90:203-222) */
static int in_obstacle(double x, double y) {
  if (g_obs.type == 1)
    return x >= g_obs.x0 - 1.0e-9 && x <= g_obs.x1 + 1.0e-9 && y >= g_obs.y0 - 1.0e-9 && y <= g_obs.y1 + 1.0e-9;
  if (g_obs.type == 2)
    return (x - g_obs.xc) * (x - g_obs.xc) + (y - g_obs.yc) * (y - g_obs.yc) <= g_obs.r * g_obs.r;
  return 0;
}

/* is_u_face_in_obstacle / is_v_face_in_obstacle (src/module_navier_stokes.f90:82-131)
 * and the obs_perm <= 1e-12 test of the Poisson block: a solid face. */
static int solid_u(const orc_mesh *m, int i, int j) {
  if (!g_obs.enabled || !(g_obs.perm <= 1.0e-12)) return 0;
  return in_obstacle(x_face(m, i), ((double)(j - 1) + 0.5) * m->dy);
}
static int solid_v(const orc_mesh *m, int i, int j) {
  if (!g_obs.enabled || !(g_obs.perm <= 1.0e-12)) return 0;
  return in_obstacle(((double)(i - 1) + 0.5) * m->dx, y_face(m, j));
}

/* src/module_simulation.f90:2477-2533 (solid obstacle faces: orc_set_obstacle). */
void orc_poisson_assemble(const orc_mesh *m, const double *u, const double *v,
                          const double *rho_u, const double *rho_v, double dt,
                          double *aW, double *aE, double *aS, double *aN, double *aP,
                          double *b) {
  const int nx = m->nx, ny = m->ny;
  const double dx = m->dx, dy = m->dy, dvol = dx * dy;
  const int bl = !per_x(m), br = !per_x(m), bb = !per_y(m), bt = !per_y(m);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const long k = W(m, i, j);
      const double div = (u[H(m, i + 1, j)] - u[H(m, i, j)]) * dy + (v[H(m, i, j + 1)] - v[H(m, i, j)]) * dx;
      b[k] = -div;
      aW[k] = aE[k] = aS[k] = aN[k] = 0.0;
      if (!((i == 1 && bl && wall(m->bc[0])) || solid_u(m, i, j))) {
        double r = rho_u[H(m, i, j)];
        if (r < 1.0e-12) r = 1.0;
        aW[k] = (dt * dy) / (r * dx);
      }
      if (!((i == nx && br && wall(m->bc[1])) || solid_u(m, i + 1, j))) {
        double r = rho_u[H(m, i + 1, j)];
        if (r < 1.0e-12) r = 1.0;
        aE[k] = (dt * dy) / (r * dx);
      }
      if (!((j == 1 && bb && wall(m->bc[2])) || solid_v(m, i, j))) {
        double r = rho_v[H(m, i, j)];
        if (r < 1.0e-12) r = 1.0;
        aS[k] = (dt * dx) / (r * dy);
      }
      if (!((j == ny && bt && wall(m->bc[3])) || solid_v(m, i, j + 1))) {
        double r = rho_v[H(m, i, j + 1)];
        if (r < 1.0e-12) r = 1.0;
        aN[k] = (dt * dx) / (r * dy);
      }
      aP[k] = aW[k] + aE[k] + aS[k] + aN[k] + 1.0e-12 * dvol;
    }
}

/* src/module_simulation.f90:2668-2707 */
void orc_poisson_matvec(const orc_mesh *m, const double *aP, const double *aW,
                        const double *aE, const double *aS, const double *aN,
                        const double *x, double *y) {
  const int nx = m->nx, ny = m->ny;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const long k = W(m, i, j);
      double val = aP[k] * x[k];
      if (i > 1) val = val - aW[k] * x[W(m, i - 1, j)];
      else if (m->bc[0] == BC_PERIODIC) val = val - aW[k] * x[W(m, nx, j)];
      if (i < nx) val = val - aE[k] * x[W(m, i + 1, j)];
      else if (m->bc[1] == BC_PERIODIC) val = val - aE[k] * x[W(m, 1, j)];
      if (j > 1) val = val - aS[k] * x[W(m, i, j - 1)];
      else if (m->bc[2] == BC_PERIODIC) val = val - aS[k] * x[W(m, i, ny)];
      if (j < ny) val = val - aN[k] * x[W(m, i, j + 1)];
      else if (m->bc[3] == BC_PERIODIC) val = val - aN[k] * x[W(m, i, 1)];
      y[k] = val;
    }
}

/* src/module_simulation.f90:2709-2758 (symmetric Gauss-Seidel, omega = 1). */
void orc_poisson_sgs(const orc_mesh *m, const double *aP, const double *aW,
                     const double *aE, const double *aS, const double *aN,
                     const double *r, double *z) {
  const int nx = m->nx, ny = m->ny;
  memset(z, 0, (size_t)nx * ny * sizeof(double));
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const long k = W(m, i, j);
      double val = r[k];
      if (i > 1) val = val + aW[k] * z[W(m, i - 1, j)];
      else if (m->bc[0] == BC_PERIODIC) val = val + aW[k] * z[W(m, nx, j)];
      if (j > 1) val = val + aS[k] * z[W(m, i, j - 1)];
      else if (m->bc[2] == BC_PERIODIC) val = val + aS[k] * z[W(m, i, ny)];
      z[k] = val / aP[k];
    }
  for (int j = ny; j >= 1; --j)
    for (int i = nx; i >= 1; --i) {
      const long k = W(m, i, j);
      double val = aP[k] * z[k];
      if (i < nx) val = val + aE[k] * z[W(m, i + 1, j)];
      else if (m->bc[1] == BC_PERIODIC) val = val + aE[k] * z[W(m, 1, j)];
      if (j < ny) val = val + aN[k] * z[W(m, i, j + 1)];
      else if (m->bc[3] == BC_PERIODIC) val = val + aN[k] * z[W(m, i, 1)];
      z[k] = val / aP[k];
    }
}

/* Jacobi: src/module_linear_solvers.f90:88-135 (z = r/d if |d| > 1e-15). */
void orc_poisson_precond(const orc_mesh *m, int precond, const double *aP, const double *aW,
                         const double *aE, const double *aS, const double *aN,
                         const double *r, double *z) {
  const long n = (long)m->nx * m->ny;
  if (precond == 2) {
    orc_poisson_sgs(m, aP, aW, aE, aS, aN, r, z);
  } else if (precond == 1) {
    for (long k = 0; k < n; ++k) z[k] = (fabs(aP[k]) > 1.0e-15) ? r[k] / aP[k] : r[k];
  } else {
    memcpy(z, r, n * sizeof(double));
  }
}

double orc_dot(const double *x, const double *y, long n) {
  double s = 0.0;
  for (long k = 0; k < n; ++k) s = s + x[k] * y[k];
  return s;
}

/* src/module_simulation.f90:2535-2575 */
int orc_sp_pcg(const orc_mesh *m, int precond, const double *aP, const double *aW,
               const double *aE, const double *aS, const double *aN,
               const double *b, double *x, double tol, int maxit, double *relres) {
  const long n = (long)m->nx * m->ny;
  double *r = malloc(n * sizeof(double)), *Ad = malloc(n * sizeof(double));
  double *z = malloc(n * sizeof(double)), *d = malloc(n * sizeof(double));
  orc_poisson_matvec(m, aP, aW, aE, aS, aN, x, Ad);
  for (long k = 0; k < n; ++k) r[k] = b[k] - Ad[k];
  double b_norm = sqrt(orc_dot(b, b, n));
  if (b_norm < 1.0e-30) b_norm = 1.0;
  double r_norm = sqrt(orc_dot(r, r, n));
  int iters = 0;
  if (r_norm / b_norm > tol) {
    orc_poisson_precond(m, precond, aP, aW, aE, aS, aN, r, z);
    memcpy(d, z, n * sizeof(double));
    double rz_old = orc_dot(r, z, n);
    for (int it = 1; it <= maxit; ++it) {
      ++iters;
      orc_poisson_matvec(m, aP, aW, aE, aS, aN, d, Ad);
      const double dAd = orc_dot(d, Ad, n);
      if (fabs(dAd) < 1.0e-35) break;
      const double alpha = rz_old / dAd;
      for (long k = 0; k < n; ++k) x[k] = x[k] + alpha * d[k];
      for (long k = 0; k < n; ++k) r[k] = r[k] - alpha * Ad[k];
      r_norm = sqrt(orc_dot(r, r, n));
      if (r_norm / b_norm <= tol) break;
      orc_poisson_precond(m, precond, aP, aW, aE, aS, aN, r, z);
      const double rz_new = orc_dot(r, z, n);
      if (fabs(rz_old) < 1.0e-35) break;
      const double beta = rz_new / rz_old;
      rz_old = rz_new;
      for (long k = 0; k < n; ++k) d[k] = z[k] + beta * d[k];
    }
  }
  if (relres) *relres = r_norm / b_norm;
  free(r); free(Ad); free(z); free(d);
  return iters;
}

/* src/module_linear_solvers.f90:361-579, with apply_PA (:323-356) = precond(A x). */
int orc_bicgstab2(const orc_mesh *m, int precond, const double *aP, const double *aW,
                  const double *aE, const double *aS, const double *aN,
                  const double *b, double *x, double tol, int maxit, double *relres) {
  const long n = (long)m->nx * m->ny;
  double *bh = malloc(n * sizeof(double)), *rt = malloc(n * sizeof(double));
  double *r0 = malloc(n * sizeof(double)), *r1 = calloc(n, sizeof(double));
  double *r2 = calloc(n, sizeof(double)), *u0 = calloc(n, sizeof(double));
  double *u1 = calloc(n, sizeof(double)), *u2 = calloc(n, sizeof(double));
  double *tmp = malloc(n * sizeof(double));
#define PA(src, dst)                                                  \
  do {                                                                \
    orc_poisson_matvec(m, aP, aW, aE, aS, aN, src, tmp);              \
    orc_poisson_precond(m, precond, aP, aW, aE, aS, aN, tmp, dst);    \
  } while (0)
  orc_poisson_precond(m, precond, aP, aW, aE, aS, aN, b, bh);
  PA(x, r0);
  for (long k = 0; k < n; ++k) r0[k] = bh[k] - r0[k];
  const double r0_norm = sqrt(orc_dot(r0, r0, n));
  int iter = 0;
  double r_norm = r0_norm;
  if (r0_norm < 1.0e-12) goto done;
  memcpy(rt, r0, n * sizeof(double));
  double rho_0 = 1.0, rho_1, alpha = 0.0, omega = 1.0, beta, gamma;
  while (iter < maxit) {
    iter += 2;
    rho_0 = -omega * rho_0;
    rho_1 = orc_dot(rt, r0, n);
    if (fabs(rho_0) < 1.0e-300) break;
    beta = alpha * (rho_1 / rho_0);
    rho_0 = rho_1;
    for (long k = 0; k < n; ++k) u0[k] = r0[k] - beta * u0[k];
    PA(u0, u1);
    gamma = orc_dot(rt, u1, n);
    if (fabs(gamma) < 1.0e-300) break;
    alpha = rho_0 / gamma;
    for (long k = 0; k < n; ++k) r0[k] = r0[k] - alpha * u1[k];
    for (long k = 0; k < n; ++k) x[k] = x[k] + alpha * u0[k];
    PA(r0, r1);
    rho_1 = orc_dot(rt, r1, n);
    if (fabs(rho_0) < 1.0e-300) break;
    beta = alpha * (rho_1 / rho_0);
    rho_0 = rho_1;
    for (long k = 0; k < n; ++k) u0[k] = r0[k] - beta * u0[k];
    for (long k = 0; k < n; ++k) u1[k] = r1[k] - beta * u1[k];
    PA(u1, u2);
    gamma = orc_dot(rt, u2, n);
    if (fabs(gamma) < 1.0e-300) break;
    alpha = rho_0 / gamma;
    for (long k = 0; k < n; ++k) x[k] = x[k] + alpha * u0[k];
    for (long k = 0; k < n; ++k) r0[k] = r0[k] - alpha * u1[k];
    for (long k = 0; k < n; ++k) r1[k] = r1[k] - alpha * u2[k];
    PA(r1, r2);
    const double s11 = orc_dot(r1, r1, n), s12 = orc_dot(r2, r1, n), s22 = orc_dot(r2, r2, n);
    const double rhs1 = orc_dot(r0, r1, n), rhs2 = orc_dot(r0, r2, n);
    const double rn1 = sqrt(s11), rn2 = sqrt(s22);
    double g1, g2;
    if (rn1 < 1.0e-300) {
      break;
    } else if (rn2 < 1.0e-300) {
      g2 = 0.0;
      g1 = rhs1 / s11;
    } else {
      const double i1 = 1.0 / rn1, i2 = 1.0 / rn2;
      const double c12 = s12 * i1 * i2, b1 = rhs1 * i1, b2 = rhs2 * i2;
      const double det = 1.0 - c12 * c12;
      if (fabs(det) < 1.0e-14) {
        g2 = 0.0;
        g1 = b1 * i1;
      } else {
        g1 = (b1 - c12 * b2) / (det * rn1);
        g2 = (b2 - c12 * b1) / (det * rn2);
      }
    }
    for (long k = 0; k < n; ++k) x[k] = x[k] + g1 * r0[k] + g2 * r1[k];
    for (long k = 0; k < n; ++k) r0[k] = r0[k] - (g1 * r1[k] + g2 * r2[k]);
    for (long k = 0; k < n; ++k) u0[k] = u0[k] - (g1 * u1[k] + g2 * u2[k]);
    if (g2 != 0.0) omega = g2;
    else if (fabs(g1) > 1.0e-300) omega = g1;
    else omega = 1.0;
    r_norm = sqrt(orc_dot(r0, r0, n));
    if (r_norm != r_norm || r_norm > 1.0e30) break;
    if (r_norm / r0_norm < tol || r_norm < 1.0e-12) break;
  }
done:
#undef PA
  if (relres) *relres = (r0_norm > 0.0) ? r_norm / r0_norm : 0.0;
  free(bh); free(rt); free(r0); free(r1); free(r2); free(u0); free(u1); free(u2); free(tmp);
  return iter;
}

/* src/module_simulation.f90:2577-2585 plus sync_cell_scalar on one rank
 * (src/module_mpi_comm.f90:46-56 semantics for the periodic wrap). */
void orc_sp_finish_pressure(const orc_mesh *m, const double *x, double *p) {
  const long n = (long)m->nx * m->ny;
  double s = 0.0;
  for (long k = 0; k < n; ++k) s = s + x[k];
  const double mean = s / (double)((long)m->nx * m->ny);
  for (int j = 1; j <= m->ny; ++j)
    for (int i = 1; i <= m->nx; ++i) p[H(m, i, j)] = x[W(m, i, j)] - mean;
  orc_exchange_halos_scalar(m, p);
}

/* src/module_simulation.f90:2590-2654 (solid obstacle faces skipped; halo/BC sync not included). */
void orc_sp_correct(const orc_mesh *m, const double *p, const double *rho_u,
                    const double *rho_v, double dt, double *u, double *v) {
  const int nx = m->nx, ny = m->ny;
  const double dx = m->dx, dy = m->dy;
  const int bl = !per_x(m), br = !per_x(m), bb = !per_y(m), bt = !per_y(m);
  const int u_nx = br ? nx + 1 : nx, v_ny = bt ? ny + 1 : ny;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= u_nx; ++i) {
      const int is_wall = (i == 1 && bl && wall(m->bc[0])) || (i == u_nx && br && wall(m->bc[1])) ||
                          solid_u(m, i, j);
      if (is_wall) continue;
      double r = rho_u[H(m, i, j)];
      if (r < 1.0e-12) r = 1.0;
      double pw, pe;
      if (i > 1) pw = p[H(m, i - 1, j)];
      else if (m->bc[0] == BC_PERIODIC) pw = p[H(m, nx, j)];
      else pw = p[H(m, 1, j)];
      if (i <= nx) pe = p[H(m, i, j)];
      else if (m->bc[1] == BC_PERIODIC) pe = p[H(m, 1, j)];
      else pe = p[H(m, nx, j)];
      const double g = (pe - pw) / dx;
      u[H(m, i, j)] = u[H(m, i, j)] - (dt / r) * g;
    }
  for (int j = 1; j <= v_ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const int is_wall = (j == 1 && bb && wall(m->bc[2])) || (j == v_ny && bt && wall(m->bc[3])) ||
                          solid_v(m, i, j);
      if (is_wall) continue;
      double r = rho_v[H(m, i, j)];
      if (r < 1.0e-12) r = 1.0;
      double ps, pn;
      if (j > 1) ps = p[H(m, i, j - 1)];
      else if (m->bc[2] == BC_PERIODIC) ps = p[H(m, i, ny)];
      else ps = p[H(m, i, 1)];
      if (j <= ny) pn = p[H(m, i, j)];
      else if (m->bc[3] == BC_PERIODIC) pn = p[H(m, i, 1)];
      else pn = p[H(m, i, ny)];
      const double g = (pn - ps) / dy;
      v[H(m, i, j)] = v[H(m, i, j)] - (dt / r) * g;
    }
}

/* CPU baseline: Jacobi-PCG iterations (same algorithm as the SP PCG loop with
 * z = r/aP), OpenMP over `threads`.  Returns wall seconds for `iters` iterations. */
double orc_bench_jacobi_pcg(const orc_mesh *m, const double *aP, const double *aW,
                            const double *aE, const double *aS, const double *aN,
                            const double *b, int iters, int threads) {
  const int nx = m->nx, ny = m->ny;
  const long n = (long)nx * ny;
  double *x = calloc(n, sizeof(double)), *r = malloc(n * sizeof(double));
  double *d = malloc(n * sizeof(double)), *q = malloc(n * sizeof(double));
  omp_set_num_threads(threads);
  double rz = 0.0;
#pragma omp parallel for reduction(+ : rz)
  for (long k = 0; k < n; ++k) {
    r[k] = b[k];
    d[k] = r[k] / aP[k];
    rz += r[k] * d[k];
  }
  const double t0 = omp_get_wtime();
  for (int it = 0; it < iters; ++it) {
    double dq = 0.0;
#pragma omp parallel for reduction(+ : dq)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const long k = W(m, i, j);
        double val = aP[k] * d[k];
        if (i > 1) val -= aW[k] * d[k - 1];
        if (i < nx) val -= aE[k] * d[k + 1];
        if (j > 1) val -= aS[k] * d[k - nx];
        if (j < ny) val -= aN[k] * d[k + nx];
        q[k] = val;
        dq += d[k] * val;
      }
    const double alpha = rz / dq;
    double rz_new = 0.0, rr = 0.0;
#pragma omp parallel for reduction(+ : rz_new, rr)
    for (long k = 0; k < n; ++k) {
      x[k] += alpha * d[k];
      r[k] -= alpha * q[k];
      rz_new += r[k] * (r[k] / aP[k]);
      rr += r[k] * r[k];
    }
    const double beta = rz_new / rz;
    rz = rz_new;
#pragma omp parallel for
    for (long k = 0; k < n; ++k) d[k] = r[k] / aP[k] + beta * d[k];
    (void)rr;
  }
  const double t1 = omp_get_wtime();
  free(x); free(r); free(d); free(q);
  return t1 - t0;
}

/* CPU baseline: the SP PCG loop with Jacobi (z = r/aP) from x = 0 to
 * ||r||/||b|| <= tol, OpenMP over `threads` (reductions in OpenMP's order).
 * Returns wall seconds; *iters and *relres as the loop ends. */
double orc_bench_jacobi_pcg_solve(const orc_mesh *m, const double *aP, const double *aW,
                                  const double *aE, const double *aS, const double *aN,
                                  const double *b, double tol, int maxit, int threads, int *iters,
                                  double *relres) {
  const int nx = m->nx, ny = m->ny;
  const long n = (long)nx * ny;
  double *x = calloc(n, sizeof(double)), *r = malloc(n * sizeof(double));
  double *d = malloc(n * sizeof(double)), *q = malloc(n * sizeof(double));
  omp_set_num_threads(threads);
  const double t0 = omp_get_wtime();
  double rz = 0.0, bb = 0.0;
#pragma omp parallel for reduction(+ : rz, bb)
  for (long k = 0; k < n; ++k) {
    r[k] = b[k];
    d[k] = r[k] / aP[k];
    rz += r[k] * d[k];
    bb += b[k] * b[k];
  }
  double bn = sqrt(bb);
  if (bn < 1.0e-30) bn = 1.0;
  double rel = 1.0;
  int it = 0;
  while (it < maxit) {
    ++it;
    double dq = 0.0;
#pragma omp parallel for reduction(+ : dq)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const long k = W(m, i, j);
        double val = aP[k] * d[k];
        if (i > 1) val -= aW[k] * d[k - 1];
        if (i < nx) val -= aE[k] * d[k + 1];
        if (j > 1) val -= aS[k] * d[k - nx];
        if (j < ny) val -= aN[k] * d[k + nx];
        q[k] = val;
        dq += d[k] * val;
      }
    const double alpha = rz / dq;
    double rz_new = 0.0, rr = 0.0;
#pragma omp parallel for reduction(+ : rz_new, rr)
    for (long k = 0; k < n; ++k) {
      x[k] += alpha * d[k];
      r[k] -= alpha * q[k];
      rz_new += r[k] * (r[k] / aP[k]);
      rr += r[k] * r[k];
    }
    rel = sqrt(rr) / bn;
    if (rel <= tol) break;
    const double beta = rz_new / rz;
    rz = rz_new;
#pragma omp parallel for
    for (long k = 0; k < n; ++k) d[k] = r[k] / aP[k] + beta * d[k];
  }
  const double t1 = omp_get_wtime();
  if (iters) *iters = it;
  if (relres) *relres = rel;
  free(x); free(r); free(d); free(q);
  return t1 - t0;
}

/* ------------------------------------------------ SP velocity predictor
 *
 * The momentum step of the standard projection (src/module_simulation.f90:
 * 2454-2469): assemble_preconditioner_blocks (diagonal), assemble_rhs,
 * pack_vector, solve_linear_system, unpack_vector, sync + velocity BC.
 * Supported subset: SOLVER_STANDARD_PROJECTION, no obstacles, UNIFORM
 * boundary profiles, no energy equation, no SSP-RK3 momentum. */

/* The reference's has_nm1 / u_old / u_nm1 inputs of the momentum step. */
static int gear(const orc_mom_cfg *c) { return c->time_scheme == 2 && c->has_nm1; }
static int ab_nm1(const orc_mom_cfg *c) { return c->conv_lin == 2 && c->has_nm1; }

/* src/module_mpi_comm.f90:290-537 (apply_velocity_bc) for UNIFORM profiles. */
void orc_velocity_bc(const orc_mesh *m, const orc_mom_cfg *c, double *u, double *v, int homogeneous) {
  const int nx = m->nx, ny = m->ny;
  const int *bc = m->bc;
  const double u_l = homogeneous ? 0.0 : c->bc_u[0], v_l = homogeneous ? 0.0 : c->bc_v[0];
  const double v_r = homogeneous ? 0.0 : c->bc_v[1];
  const double u_r = homogeneous ? 0.0 : c->bc_u[1];
  const double u_b = homogeneous ? 0.0 : c->bc_u[2], v_b = homogeneous ? 0.0 : c->bc_v[2];
  const double u_t = homogeneous ? 0.0 : c->bc_u[3], v_t = homogeneous ? 0.0 : c->bc_v[3];
  if (bc[0] == BC_PERIODIC) {
    for (int j = 1; j <= ny; ++j) {
      u[H(m, 0, j)] = u[H(m, nx, j)];
      u[H(m, nx + 1, j)] = u[H(m, 1, j)];
      v[H(m, 0, j)] = v[H(m, nx, j)];
      v[H(m, nx + 1, j)] = v[H(m, 1, j)];
    }
  }
  if (bc[2] == BC_PERIODIC) {
    for (int i = 1; i <= nx; ++i) {
      v[H(m, i, 0)] = v[H(m, i, ny)];
      v[H(m, i, ny + 1)] = v[H(m, i, 1)];
    }
    for (int i = 0; i <= nx + 1; ++i) {
      u[H(m, i, 0)] = u[H(m, i, ny)];
      u[H(m, i, ny + 1)] = u[H(m, i, 1)];
    }
  }
  if (!per_x(m) && bc[0] != BC_PERIODIC) {
    for (int j = 1; j <= ny; ++j) {
      if (bc[0] == BC_DIRICHLET) {
        const double u_inlet = u_l;
        if (!homogeneous) u[H(m, 1, j)] = u_inlet;
        u[H(m, 0, j)] = u_inlet;
        v[H(m, 0, j)] = 2.0 * v_l - v[H(m, 1, j)];
      } else if (bc[0] == BC_FREESLIP) {
        if (!homogeneous) u[H(m, 1, j)] = 0.0;
        u[H(m, 0, j)] = 0.0;
        v[H(m, 0, j)] = v[H(m, 1, j)];
      } else if (bc[0] == BC_NEUMANN) {
        u[H(m, 0, j)] = u[H(m, 1, j)];
        v[H(m, 0, j)] = v[H(m, 1, j)];
      }
    }
  }
  if (!per_x(m) && bc[1] != BC_PERIODIC) {
    for (int j = 1; j <= ny; ++j) {
      if (bc[1] == BC_DIRICHLET) {
        if (!homogeneous) u[H(m, nx + 1, j)] = u_r;
        v[H(m, nx + 1, j)] = 2.0 * v_r - v[H(m, nx, j)];
      } else if (bc[1] == BC_FREESLIP) {
        if (!homogeneous) u[H(m, nx + 1, j)] = 0.0;
        v[H(m, nx + 1, j)] = v[H(m, nx, j)];
      } else if (bc[1] == BC_NEUMANN) {
        if (!homogeneous) u[H(m, nx + 1, j)] = u[H(m, nx, j)];
        v[H(m, nx + 1, j)] = v[H(m, nx, j)];
      }
    }
  }
  if (!per_y(m) && bc[2] != BC_PERIODIC) {
    for (int i = 1; i <= nx; ++i) {
      if (bc[2] == BC_DIRICHLET) {
        if (!homogeneous) v[H(m, i, 1)] = v_b;
        v[H(m, i, 0)] = v_b;
        if (ny >= 2) u[H(m, i, 0)] = (8.0 * u_b - 6.0 * u[H(m, i, 1)] + u[H(m, i, 2)]) / 3.0;
        else u[H(m, i, 0)] = 2.0 * u_b - u[H(m, i, 1)];
      } else if (bc[2] == BC_FREESLIP) {
        if (!homogeneous) v[H(m, i, 1)] = 0.0;
        v[H(m, i, 0)] = 0.0;
        u[H(m, i, 0)] = u[H(m, i, 1)];
      } else if (bc[2] == BC_NEUMANN) {
        u[H(m, i, 0)] = u[H(m, i, 1)];
        v[H(m, i, 0)] = v[H(m, i, 1)];
      }
    }
  }
  if (!per_y(m) && bc[3] != BC_PERIODIC) {
    for (int i = 1; i <= nx; ++i) {
      if (bc[3] == BC_DIRICHLET) {
        if (!homogeneous) v[H(m, i, ny + 1)] = v_t;
        if (ny >= 2) u[H(m, i, ny + 1)] = (8.0 * u_t - 6.0 * u[H(m, i, ny)] + u[H(m, i, ny - 1)]) / 3.0;
        else u[H(m, i, ny + 1)] = 2.0 * u_t - u[H(m, i, ny)];
      } else if (bc[3] == BC_FREESLIP) {
        if (!homogeneous) v[H(m, i, ny + 1)] = 0.0;
        u[H(m, i, ny + 1)] = u[H(m, i, ny)];
      } else if (bc[3] == BC_NEUMANN) {
        u[H(m, i, ny + 1)] = u[H(m, i, ny)];
        if (!homogeneous) v[H(m, i, ny + 1)] = v[H(m, i, ny)];
      }
    }
  }
  if (!homogeneous) {
    if (!per_x(m) && bc[1] != BC_PERIODIC) {
      if (!per_y(m) && bc[2] != BC_PERIODIC) {
        u[H(m, nx + 1, 0)] = u[H(m, nx + 1, 1)];
        v[H(m, nx + 1, 0)] = v[H(m, nx, 0)];
      }
      if (!per_y(m) && bc[3] != BC_PERIODIC) {
        u[H(m, nx + 1, ny + 1)] = u[H(m, nx + 1, ny)];
        v[H(m, nx + 1, ny + 1)] = v[H(m, nx, ny + 1)];
      }
    }
    if (!per_x(m) && bc[0] != BC_PERIODIC && !per_y(m) && bc[3] != BC_PERIODIC)
      v[H(m, 0, ny + 1)] = v[H(m, 1, ny + 1)];
  }
}

/* Unknown counts of ns_init (src/module_navier_stokes.f90:136-226), SP. */
void orc_mom_sizes(const orc_mesh *m, int *u_nx, int *v_ny) {
  *u_nx = m->nx + (per_x(m) ? 0 : 1);
  *v_ny = m->ny + (per_y(m) ? 0 : 1);
}

/* ns_pack_vector (src/module_navier_stokes.f90:231-295), SP unknowns. */
void orc_mom_pack(const orc_mesh *m, const double *u, const double *v, double *vec) {
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const long nu = (long)u_nx * m->ny;
  for (int j = 1; j <= m->ny; ++j)
    for (int i = 1; i <= u_nx; ++i) vec[(long)(j - 1) * u_nx + (i - 1)] = u[H(m, i, j)];
  for (int j = 1; j <= v_ny; ++j)
    for (int i = 1; i <= m->nx; ++i) vec[nu + (long)(j - 1) * m->nx + (i - 1)] = v[H(m, i, j)];
}

/* ns_unpack_vector (src/module_navier_stokes.f90:300-375): interior unknowns,
 * sync_u / sync_v (the single-rank exchange), apply_boundary_conditions. */
void orc_mom_unpack(const orc_mesh *m, const orc_mom_cfg *c, const double *vec, double *u, double *v,
                    int homogeneous) {
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const long nu = (long)u_nx * m->ny;
  for (int j = 1; j <= m->ny; ++j)
    for (int i = 1; i <= u_nx; ++i) u[H(m, i, j)] = vec[(long)(j - 1) * u_nx + (i - 1)];
  for (int j = 1; j <= v_ny; ++j)
    for (int i = 1; i <= m->nx; ++i) v[H(m, i, j)] = vec[nu + (long)(j - 1) * m->nx + (i - 1)];
  orc_exchange_halos_scalar(m, u);
  orc_exchange_halos_scalar(m, v);
  orc_velocity_bc(m, c, u, v, homogeneous);
}

/* conv_face_flux (src/module_navier_stokes.f90:2388-2416). */
static double conv_face_flux(int scheme, double flux, double phi_L, double phi_R, double q_LL,
                             double q_L, double q_R, double q_RR, int has_LL, int has_RR) {
  if (scheme == 2) return 0.5 * flux * (phi_L + phi_R);
  if (scheme == 1) return flux >= 0.0 ? flux * phi_L : flux * phi_R;
  const double dq = q_R - q_L;
  double psi = 0.0, r;
  if (flux >= 0.0) {
    if (has_LL && fabs(dq) > 1.0e-14) {
      r = (q_L - q_LL) / dq;
      psi = dmax(dmax(0.0, dmin(2.0 * r, 1.0)), dmin(r, 2.0));
    }
    return flux * (phi_L + 0.5 * psi * (phi_R - phi_L));
  }
  if (has_RR && fabs(dq) > 1.0e-14) {
    r = (q_RR - q_R) / dq;
    psi = dmax(dmax(0.0, dmin(2.0 * r, 1.0)), dmin(r, 2.0));
  }
  return flux * (phi_R + 0.5 * psi * (phi_L - phi_R));
}

/* compute_u_conv_fluxes (src/module_navier_stokes.f90:2418-2457). */
static void u_conv_fluxes(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, int i,
                          int j, double *FE, double *FW, double *FN, double *FS) {
  const double *uu = c->conv_lin == 2 ? F->u_old : F->u, *vv = c->conv_lin == 2 ? F->v_old : F->v;
  double u_n = uu[H(m, i, j)], u_n_E = uu[H(m, i + 1, j)], u_n_W = uu[H(m, i - 1, j)];
  double v_n_N = 0.5 * (vv[H(m, i - 1, j + 1)] + vv[H(m, i, j + 1)]);
  double v_n_S = 0.5 * (vv[H(m, i - 1, j)] + vv[H(m, i, j)]);
  if (ab_nm1(c)) {
    u_n = 2.0 * u_n - F->u_nm1[H(m, i, j)];
    u_n_E = 2.0 * u_n_E - F->u_nm1[H(m, i + 1, j)];
    u_n_W = 2.0 * u_n_W - F->u_nm1[H(m, i - 1, j)];
    const double v_nm1_N = 0.5 * (F->v_nm1[H(m, i - 1, j + 1)] + F->v_nm1[H(m, i, j + 1)]);
    const double v_nm1_S = 0.5 * (F->v_nm1[H(m, i - 1, j)] + F->v_nm1[H(m, i, j)]);
    v_n_N = 2.0 * v_n_N - v_nm1_N;
    v_n_S = 2.0 * v_n_S - v_nm1_S;
  }
  *FE = 0.5 * (u_n + u_n_E) * m->dy;
  *FW = 0.5 * (u_n_W + u_n) * m->dy;
  *FN = v_n_N * m->dx;
  *FS = v_n_S * m->dx;
}

/* compute_v_conv_fluxes (src/module_navier_stokes.f90:2459-2498). */
static void v_conv_fluxes(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, int i,
                          int j, double *FE, double *FW, double *FN, double *FS) {
  const double *uu = c->conv_lin == 2 ? F->u_old : F->u, *vv = c->conv_lin == 2 ? F->v_old : F->v;
  double v_n = vv[H(m, i, j)], v_n_N = vv[H(m, i, j + 1)], v_n_S = vv[H(m, i, j - 1)];
  double u_n_E = 0.5 * (uu[H(m, i + 1, j - 1)] + uu[H(m, i + 1, j)]);
  double u_n_W = 0.5 * (uu[H(m, i, j - 1)] + uu[H(m, i, j)]);
  if (ab_nm1(c)) {
    v_n = 2.0 * v_n - F->v_nm1[H(m, i, j)];
    v_n_N = 2.0 * v_n_N - F->v_nm1[H(m, i, j + 1)];
    v_n_S = 2.0 * v_n_S - F->v_nm1[H(m, i, j - 1)];
    const double u_nm1_E = 0.5 * (F->u_nm1[H(m, i + 1, j - 1)] + F->u_nm1[H(m, i + 1, j)]);
    const double u_nm1_W = 0.5 * (F->u_nm1[H(m, i, j - 1)] + F->u_nm1[H(m, i, j)]);
    u_n_E = 2.0 * u_n_E - u_nm1_E;
    u_n_W = 2.0 * u_n_W - u_nm1_W;
  }
  *FE = u_n_E * m->dy;
  *FW = u_n_W * m->dy;
  *FN = 0.5 * (v_n + v_n_N) * m->dx;
  *FS = 0.5 * (v_n_S + v_n) * m->dx;
}

/* ns_apply_system_matrix (src/module_navier_stokes.f90:380-780), SP branch:
 * tu, tv are the persistent work fields of the reference's `tf` (SAVE),
 * which unpack_vector(homogeneous) fills before every product; the caller
 * keeps them (zero-initialised, like the reference's first allocation). */
/* The u and v rows of the operator from unpacked work fields; with tp (the
 * monolithic system) the pressure-gradient terms B_u^T p, B_v^T p too. */
static void mom_rows(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                     const double *tu, const double *tv, const double *tp, double *y) {
  const int nx = m->nx, ny = m->ny;
  const double dx = m->dx, dy = m->dy, dvol = dx * dy;
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const int bl = !per_x(m), br = !per_x(m), bb = !per_y(m), bt = !per_y(m);
  const long nu = (long)u_nx * ny;
  const double at = gear(c) ? 1.5 : 1.0;
  const int sch = c->conv_scheme;
  const double *u = F->u, *v = F->v;
#define TU(a, b) tu[H(m, a, b)]
#define TV(a, b) tv[H(m, a, b)]
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= u_nx; ++i) {
      const long dof = (long)(j - 1) * u_nx + (i - 1);
      const double rho_u = F->rho_u[H(m, i, j)];
      const double diag_bc = (at * rho_u / dt) * dvol;
      if (i == 1 && bl) {
        y[dof] = m->bc[0] == BC_NEUMANN ? diag_bc * (TU(1, j) - TU(2, j)) : diag_bc * TU(1, j);
        continue;
      }
      if (i == u_nx && br) {
        y[dof] = m->bc[1] == BC_NEUMANN ? diag_bc * (TU(u_nx, j) - TU(u_nx - 1, j)) : diag_bc * TU(u_nx, j);
        continue;
      }
      double val = (at * rho_u / dt) * TU(i, j) * dvol;
      if (c->enable_convection) {
        double cE, cW, cN, cS;
        u_conv_fluxes(m, c, F, i, j, &cE, &cW, &cN, &cS);
        const int has_ee = i + 2 <= nx + 1, has_ww = i - 2 >= 0, has_nn = j + 2 <= ny + 1, has_ss = j - 2 >= 0;
        const double q_ee = has_ee ? u[H(m, i + 2, j)] : 0.0, q_ww = has_ww ? u[H(m, i - 2, j)] : 0.0;
        const double q_nn = has_nn ? u[H(m, i, j + 2)] : 0.0, q_ss = has_ss ? u[H(m, i, j - 2)] : 0.0;
        const double pE = conv_face_flux(sch, cE, TU(i, j), TU(i + 1, j), u[H(m, i - 1, j)], u[H(m, i, j)],
                                         u[H(m, i + 1, j)], q_ee, 1, has_ee);
        const double pW = conv_face_flux(sch, cW, TU(i - 1, j), TU(i, j), q_ww, u[H(m, i - 1, j)], u[H(m, i, j)],
                                         u[H(m, i + 1, j)], has_ww, 1);
        const double pN = conv_face_flux(sch, cN, TU(i, j), TU(i, j + 1), u[H(m, i, j - 1)], u[H(m, i, j)],
                                         u[H(m, i, j + 1)], q_nn, 1, has_nn);
        const double pS = conv_face_flux(sch, cS, TU(i, j - 1), TU(i, j), q_ss, u[H(m, i, j - 1)], u[H(m, i, j)],
                                         u[H(m, i, j + 1)], has_ss, 1);
        val = val + rho_u * (pE - pW + pN - pS);
        if (c->conv_lin == 3) {
          const double du_dx = (u[H(m, i + 1, j)] - u[H(m, i - 1, j)]) / (2.0 * dx);
          const double du_dy = (u[H(m, i, j + 1)] - u[H(m, i, j - 1)]) / (2.0 * dy);
          const double tvf = 0.25 * (TV(i - 1, j) + TV(i, j) + TV(i - 1, j + 1) + TV(i, j + 1));
          val = val + rho_u * (TU(i, j) * du_dx + tvf * du_dy) * dvol;
        }
      }
      const double mu_E = F->mu[H(m, i, j)], mu_W = F->mu[H(m, i - 1, j)];
      const double deE = 2.0 * mu_E * (TU(i + 1, j) - TU(i, j)) / dx * dy;
      const double deW = 2.0 * mu_W * (TU(i, j) - TU(i - 1, j)) / dx * dy;
      val = val - (deE - deW);
      const double mnN = F->mu_node[H(m, i, j + 1)], mnS = F->mu_node[H(m, i, j)];
      const double shN = mnN * ((TU(i, j + 1) - TU(i, j)) / dy + (TV(i, j + 1) - TV(i - 1, j + 1)) / dx) * dx;
      const double shS = mnS * ((TU(i, j) - TU(i, j - 1)) / dy + (TV(i, j) - TV(i - 1, j)) / dx) * dx;
      val = val - (shN - shS);
      if (tp) val = val + (tp[H(m, i, j)] - tp[H(m, i - 1, j)]) * dy;
      y[dof] = val;
    }
  for (int j = 1; j <= v_ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const long dof = nu + (long)(j - 1) * nx + (i - 1);
      const double rho_v = F->rho_v[H(m, i, j)];
      const double diag_bc = (at * rho_v / dt) * dvol;
      const int wb = j == 1 && bb && wall(m->bc[2]), wt = j == v_ny && bt && wall(m->bc[3]);
      if (wb || wt) {
        y[dof] = diag_bc * TV(i, j);
        continue;
      }
      double val = (at * rho_v / dt) * TV(i, j) * dvol;
      if (c->enable_convection) {
        double cE, cW, cN, cS;
        v_conv_fluxes(m, c, F, i, j, &cE, &cW, &cN, &cS);
        const int has_ee = i + 2 <= nx + 1, has_ww = i - 2 >= 0, has_nn = j + 2 <= ny + 1, has_ss = j - 2 >= 0;
        const double q_ee = has_ee ? v[H(m, i + 2, j)] : 0.0, q_ww = has_ww ? v[H(m, i - 2, j)] : 0.0;
        const double q_nn = has_nn ? v[H(m, i, j + 2)] : 0.0, q_ss = has_ss ? v[H(m, i, j - 2)] : 0.0;
        const double pE = conv_face_flux(sch, cE, TV(i, j), TV(i + 1, j), v[H(m, i - 1, j)], v[H(m, i, j)],
                                         v[H(m, i + 1, j)], q_ee, 1, has_ee);
        const double pW = conv_face_flux(sch, cW, TV(i - 1, j), TV(i, j), q_ww, v[H(m, i - 1, j)], v[H(m, i, j)],
                                         v[H(m, i + 1, j)], has_ww, 1);
        const double pN = conv_face_flux(sch, cN, TV(i, j), TV(i, j + 1), v[H(m, i, j - 1)], v[H(m, i, j)],
                                         v[H(m, i, j + 1)], q_nn, 1, has_nn);
        const double pS = conv_face_flux(sch, cS, TV(i, j - 1), TV(i, j), q_ss, v[H(m, i, j - 1)], v[H(m, i, j)],
                                         v[H(m, i, j + 1)], has_ss, 1);
        val = val + rho_v * (pE - pW + pN - pS);
        if (c->conv_lin == 3) {
          const double dv_dx = (v[H(m, i + 1, j)] - v[H(m, i - 1, j)]) / (2.0 * dx);
          const double dv_dy = (v[H(m, i, j + 1)] - v[H(m, i, j - 1)]) / (2.0 * dy);
          const double tuf = 0.25 * (TU(i, j - 1) + TU(i + 1, j - 1) + TU(i, j) + TU(i + 1, j));
          val = val + rho_v * (tuf * dv_dx + TV(i, j) * dv_dy) * dvol;
        }
      }
      const double mu_N = F->mu[H(m, i, j)], mu_S = F->mu[H(m, i, j - 1)];
      const double deN = 2.0 * mu_N * (TV(i, j + 1) - TV(i, j)) / dy * dx;
      const double deS = 2.0 * mu_S * (TV(i, j) - TV(i, j - 1)) / dy * dx;
      val = val - (deN - deS);
      const double mnE = F->mu_node[H(m, i + 1, j)], mnW = F->mu_node[H(m, i, j)];
      const double shE = mnE * ((TV(i + 1, j) - TV(i, j)) / dx + (TU(i + 1, j) - TU(i + 1, j - 1)) / dy) * dy;
      const double shW = mnW * ((TV(i, j) - TV(i - 1, j)) / dx + (TU(i, j) - TU(i, j - 1)) / dy) * dy;
      val = val - (shE - shW);
      if (tp) val = val + (tp[H(m, i, j)] - tp[H(m, i, j - 1)]) * dx;
      y[dof] = val;
    }
#undef TU
#undef TV
}

void orc_mom_matvec(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                    double *tu, double *tv, const double *x, double *y) {
  orc_mom_unpack(m, c, x, tu, tv, 1);
  mom_rows(m, c, F, dt, tu, tv, NULL, y);
}

/* ns_unpack_vector of the monolithic system: u, v, then p with its
 * sync_cell_scalar, then the u, v exchange and velocity BC. */
void orc_mono_unpack(const orc_mesh *m, const orc_mom_cfg *c, const double *vec, double *u, double *v,
                     double *p, int homogeneous) {
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const long off = (long)u_nx * m->ny + (long)m->nx * v_ny;
  for (int j = 1; j <= m->ny; ++j)
    for (int i = 1; i <= m->nx; ++i) p[H(m, i, j)] = vec[off + W(m, i, j)];
  orc_exchange_halos_scalar(m, p);
  orc_mom_unpack(m, c, vec, u, v, homogeneous);
}

/* ns_pack_vector of the monolithic system. */
void orc_mono_pack(const orc_mesh *m, const double *u, const double *v, const double *p, double *vec) {
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const long off = (long)u_nx * m->ny + (long)m->nx * v_ny;
  orc_mom_pack(m, u, v, vec);
  for (int j = 1; j <= m->ny; ++j)
    for (int i = 1; i <= m->nx; ++i) vec[off + W(m, i, j)] = p[H(m, i, j)];
}

/* ns_apply_system_matrix, SOLVER_MONOLITHIC (src/module_navier_stokes.f90:
 * 380-780): the momentum rows with B^T p, and the continuity rows
 * div(u) dvol (+ 1e-10 p(1,1) dvol at cell (1,1), incompressible). */
void orc_mono_matvec(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                     double *tu, double *tv, double *tp, const double *x, double *y) {
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const long off = (long)u_nx * m->ny + (long)m->nx * v_ny;
  const double dx = m->dx, dy = m->dy, dvol = dx * dy;
  orc_mono_unpack(m, c, x, tu, tv, tp, 1);
  mom_rows(m, c, F, dt, tu, tv, tp, y);
  for (int j = 1; j <= m->ny; ++j)
    for (int i = 1; i <= m->nx; ++i) {
      const double div = (tu[H(m, i + 1, j)] - tu[H(m, i, j)]) / dx + (tv[H(m, i, j + 1)] - tv[H(m, i, j)]) / dy;
      y[off + W(m, i, j)] = (i == 1 && j == 1) ? div * dvol + 1.0e-10 * tp[H(m, 1, 1)] * dvol : div * dvol;
    }
}

/* ns_assemble_preconditioner_blocks (src/module_navier_stokes.f90:1268-1595):
 * the diagonals diag_u, diag_v (and diag_p = schur_aP, monolithic), the
 * momentum SSOR coefficients u_a*, v_a* and the Schur coefficients schur_a*
 * (the latter (nx, ny), zero except in the monolithic system).  Any output
 * pointer of B may be NULL. */
void orc_mom_blocks(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                    const orc_mom_blocks_t *B) {
  const int nx = m->nx, ny = m->ny;
  const double dx = m->dx, dy = m->dy, dvol = dx * dy;
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const int bl = !per_x(m), br = !per_x(m), bb = !per_y(m), bt = !per_y(m);
  const long nu = (long)u_nx * ny, nv = (long)nx * v_ny;
  const double dt_safe = dt > 1.0e-30 ? dt : 1.0;
  const double at = gear(c) ? 1.5 : 1.0;
  const int mono = c->solver_type == 2;
  const long nh = (long)(nx + 2) * (ny + 2);
  double *uaP2 = malloc(nh * sizeof(double)), *vaP2 = malloc(nh * sizeof(double));
  for (long k = 0; k < nh; ++k) uaP2[k] = vaP2[k] = 1.0;
#define SETU(arr, val) do { if (B->arr) B->arr[dof] = (val); } while (0)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= u_nx; ++i) {
      const long dof = (long)(j - 1) * u_nx + (i - 1);
      const double rho_u = F->rho_u[H(m, i, j)];
      double Du = (at * rho_u / dt_safe) * dvol;
      if (Du <= 1.0e-30) Du = 1.0;
      if (i == 1 && bl) {
        SETU(diag, Du); uaP2[H(m, i, j)] = Du;
        SETU(uW, 0.0); SETU(uN, 0.0); SETU(uS, 0.0);
        SETU(uE, (m->bc[0] == BC_NEUMANN && u_nx >= 2) ? Du : 0.0);
        SETU(uP, Du);
        continue;
      }
      if (i == u_nx && br) {
        SETU(diag, Du); uaP2[H(m, i, j)] = Du;
        SETU(uE, 0.0); SETU(uN, 0.0); SETU(uS, 0.0);
        SETU(uW, (m->bc[1] == BC_NEUMANN && u_nx >= 2) ? Du : 0.0);
        SETU(uP, Du);
        continue;
      }
      const double mu_e = F->mu[H(m, i, j)], mu_w = F->mu[H(m, i - 1, j)];
      const double cE = 2.0 * mu_e * (dy / dx), cW = 2.0 * mu_w * (dy / dx);
      const double cN_raw = F->mu_node[H(m, i, j + 1)] * (dx / dy), cS_raw = F->mu_node[H(m, i, j)] * (dx / dy);
      double cN_visc = cN_raw, cS_visc = cS_raw, dwy = 0.0, cN_extra = 0.0, cS_extra = 0.0;
      if (j == 1 && bb) {
        if (m->bc[2] == BC_DIRICHLET) {
          if (ny >= 2) {
            dwy = dwy + 3.0 * cS_raw;
            cN_extra = cN_extra + cS_raw / 3.0;
          } else {
            dwy = dwy + 2.0 * cS_raw;
          }
        }
        cS_visc = 0.0;
      }
      if (j == ny && bt) {
        if (m->bc[3] == BC_DIRICHLET) {
          if (ny >= 2) {
            dwy = dwy + 3.0 * cN_raw;
            cS_extra = cS_extra + cN_raw / 3.0;
          } else {
            dwy = dwy + 2.0 * cN_raw;
          }
        }
        cN_visc = 0.0;
      }
      double conv_diag = 0.0, conv_W = 0.0, conv_E = 0.0, conv_S = 0.0, conv_N = 0.0, newt = 0.0;
      if (c->enable_convection) {
        double FE, FW, FN, FS;
        u_conv_fluxes(m, c, F, i, j, &FE, &FW, &FN, &FS);
        conv_diag = rho_u * (dmax(FE, 0.0) + dmax(-FW, 0.0) + dmax(FN, 0.0) + dmax(-FS, 0.0));
        conv_E = rho_u * dmax(-FE, 0.0);
        conv_W = rho_u * dmax(FW, 0.0);
        conv_N = (j == ny && bt) ? 0.0 : rho_u * dmax(-FN, 0.0);
        conv_S = (j == 1 && bb) ? 0.0 : rho_u * dmax(FS, 0.0);
        if (c->conv_lin == 3) {
          const double du_dx = (F->u[H(m, i + 1, j)] - F->u[H(m, i - 1, j)]) / (2.0 * dx);
          newt = rho_u * dmax(0.0, du_dx) * dvol;
        }
      }
      SETU(uW, cW + conv_W + 0.0);
      SETU(uE, cE + conv_E + 0.0);
      SETU(uS, cS_visc + conv_S + cS_extra);
      SETU(uN, cN_visc + conv_N + cN_extra);
      double aP = Du + dwy + conv_diag + 0.0 + newt + cW + cE + cS_visc + cN_visc + 0.0 + 0.0;
      if (aP <= 1.0e-15) aP = 1.0;
      SETU(diag, aP); SETU(uP, aP); uaP2[H(m, i, j)] = aP;
    }
  orc_exchange_halos_scalar(m, uaP2);
#undef SETU
#define SETV(arr, val) do { if (B->arr) B->arr[dof] = (val); } while (0)
  for (int j = 1; j <= v_ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const long dof = (long)(j - 1) * nx + (i - 1);
      const double rho_v = F->rho_v[H(m, i, j)];
      double Dv = (at * rho_v / dt_safe) * dvol;
      if (Dv <= 1.0e-30) Dv = 1.0;
      const int is_wall = (j == 1 && bb && wall(m->bc[2])) || (j == v_ny && bt && wall(m->bc[3])) ||
                          solid_v(m, i, j);
      if (is_wall) {
        if (B->diag) B->diag[nu + dof] = Dv;
        vaP2[H(m, i, j)] = Dv;
        SETV(vE, 0.0); SETV(vW, 0.0); SETV(vN, 0.0); SETV(vS, 0.0); SETV(vP, Dv);
        continue;
      }
      if (j == 1 && bb && m->bc[2] == BC_NEUMANN && v_ny >= 2) {
        if (B->diag) B->diag[nu + dof] = Dv;
        vaP2[H(m, i, j)] = Dv;
        SETV(vE, 0.0); SETV(vW, 0.0); SETV(vS, 0.0); SETV(vN, Dv); SETV(vP, Dv);
        continue;
      }
      if (j == v_ny && bt && m->bc[3] == BC_NEUMANN && v_ny >= 2) {
        if (B->diag) B->diag[nu + dof] = Dv;
        vaP2[H(m, i, j)] = Dv;
        SETV(vE, 0.0); SETV(vW, 0.0); SETV(vN, 0.0); SETV(vS, Dv); SETV(vP, Dv);
        continue;
      }
      const double mu_n = F->mu[H(m, i, j)], mu_s = F->mu[H(m, i, j - 1)];
      const double cN = 2.0 * mu_n * (dx / dy), cS = 2.0 * mu_s * (dx / dy);
      const double cE_raw = F->mu_node[H(m, i + 1, j)] * (dy / dx), cW_raw = F->mu_node[H(m, i, j)] * (dy / dx);
      double cE_visc = cE_raw, cW_visc = cW_raw, dwx = 0.0;
      if (i == 1 && bl) {
        if (m->bc[0] == BC_DIRICHLET) dwx = dwx + 2.0 * cW_raw;
        cW_visc = 0.0;
      }
      if (i == nx && br) {
        if (m->bc[1] == BC_DIRICHLET) dwx = dwx + 2.0 * cE_raw;
        cE_visc = 0.0;
      }
      double conv_diag = 0.0, conv_W = 0.0, conv_E = 0.0, conv_S = 0.0, conv_N = 0.0, newt = 0.0;
      if (c->enable_convection) {
        double FE, FW, FN, FS;
        v_conv_fluxes(m, c, F, i, j, &FE, &FW, &FN, &FS);
        conv_diag = rho_v * (dmax(FE, 0.0) + dmax(-FW, 0.0) + dmax(FN, 0.0) + dmax(-FS, 0.0));
        conv_E = (i == nx && br) ? 0.0 : rho_v * dmax(-FE, 0.0);
        conv_W = (i == 1 && bl) ? 0.0 : rho_v * dmax(FW, 0.0);
        conv_N = rho_v * dmax(-FN, 0.0);
        conv_S = rho_v * dmax(FS, 0.0);
        if (c->conv_lin == 3) {
          const double dv_dy = (F->v[H(m, i, j + 1)] - F->v[H(m, i, j - 1)]) / (2.0 * dy);
          newt = rho_v * dmax(0.0, dv_dy) * dvol;
        }
      }
      SETV(vW, cW_visc + conv_W);
      SETV(vE, cE_visc + conv_E);
      SETV(vS, cS + conv_S + 0.0);
      SETV(vN, cN + conv_N + 0.0);
      double aP = Dv + dwx + conv_diag + 0.0 + newt + cW_visc + cE_visc + cS + cN + 0.0 + 0.0;
      if (aP <= 1.0e-15) aP = 1.0;
      if (B->diag) B->diag[nu + dof] = aP;
      SETV(vP, aP);
      vaP2[H(m, i, j)] = aP;
    }
  orc_exchange_halos_scalar(m, vaP2);
#undef SETV
  if (mono) {
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const long k = W(m, i, j);
        double aW = 0.0, aE = 0.0, aS = 0.0, aN = 0.0;
        if (!(i == 1 && bl && wall(m->bc[0]))) aW = (dy * dy) / dmax(1.0e-30, uaP2[H(m, i, j)]);
        if (!(i == nx && br && wall(m->bc[1]))) aE = (dy * dy) / dmax(1.0e-30, uaP2[H(m, i + 1, j)]);
        if (!(j == 1 && bb && wall(m->bc[2]))) aS = (dx * dx) / dmax(1.0e-30, vaP2[H(m, i, j)]);
        if (!(j == ny && bt && wall(m->bc[3]))) aN = (dx * dx) / dmax(1.0e-30, vaP2[H(m, i, j + 1)]);
        const double reg = (i == 1 && j == 1) ? 1.0e-10 * dvol : 1.0e-12 * dvol;
        const double aP = aW + aE + aS + aN + 0.0 + reg;
        if (B->sW) { B->sW[k] = aW; B->sE[k] = aE; B->sS[k] = aS; B->sN[k] = aN; B->sP[k] = aP; }
        if (B->diag) B->diag[nu + nv + k] = aP;
      }
  } else if (B->sW) {
    for (long k = 0; k < (long)nx * ny; ++k) {
      B->sW[k] = B->sE[k] = B->sS[k] = B->sN[k] = 0.0;
      B->sP[k] = 1.0;
    }
  }
  free(uaP2);
  free(vaP2);
}

void orc_mom_diag(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                  double *diag) {
  orc_mom_blocks_t B;
  memset(&B, 0, sizeof(B));
  B.diag = diag;
  orc_mom_blocks(m, c, F, dt, &B);
}

/* ssor_solve_u / _v / _p (src/module_linear_solvers.f90:1278-1403): n_sweeps
 * symmetric lexicographic Gauss-Seidel sweeps with relaxation omega from
 * zero on a (ni, nj) block of unknowns (dof (j-1) ni + i-1) held in the halo
 * field x (0:nx+1, 0:ny+1), aP / a_nb the block's coefficients; with sync
 * (a periodic side on one rank) the single-rank exchange of x before every
 * sweep but the first and between the two halves of a sweep. */
void orc_ssor_block(const orc_mesh *m, int ni, int nj, const double *aW, const double *aE, const double *aS,
                    const double *aN, const double *aP, const double *rhs, double *x, int n_sweeps,
                    double omega, int sync) {
  const long nh = (long)(m->nx + 2) * (m->ny + 2);
  for (long k = 0; k < nh; ++k) x[k] = 0.0;
#define GS_CELL                                                              \
  do {                                                                       \
    const long dof = (long)(j - 1) * ni + (i - 1);                           \
    double d = aP[dof];                                                      \
    if (fabs(d) < 1.0e-30) d = 1.0;                                          \
    double val = rhs[dof];                                                   \
    if (aW[dof] != 0.0) val = val + aW[dof] * x[H(m, i - 1, j)];             \
    if (aE[dof] != 0.0) val = val + aE[dof] * x[H(m, i + 1, j)];             \
    if (aS[dof] != 0.0) val = val + aS[dof] * x[H(m, i, j - 1)];             \
    if (aN[dof] != 0.0) val = val + aN[dof] * x[H(m, i, j + 1)];             \
    x[H(m, i, j)] = (1.0 - omega) * x[H(m, i, j)] + omega * (val / d);       \
  } while (0)
  for (int sweep = 1; sweep <= n_sweeps; ++sweep) {
    if (sync && sweep > 1) orc_exchange_halos_scalar(m, x);
    for (int j = 1; j <= nj; ++j)
      for (int i = 1; i <= ni; ++i) GS_CELL;
    if (sync) orc_exchange_halos_scalar(m, x);
    for (int j = nj; j >= 1; --j)
      for (int i = ni; i >= 1; --i) GS_CELL;
  }
#undef GS_CELL
}

/* apply_preconditioner (src/module_linear_solvers.f90:48-316) for NONE,
 * JACOBI, SSOR and BLOCK_TRIANGULAR on the SP or monolithic unknowns; B holds
 * the blocks of orc_mom_blocks. */
void orc_mom_precond(const orc_mesh *m, const orc_mom_cfg *c, int precond, const orc_mom_blocks_t *B,
                     const double *src, double *dst) {
  const int nx = m->nx, ny = m->ny;
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const long nu = (long)u_nx * ny, nv = (long)nx * v_ny, np = (long)nx * ny;
  const int mono = c->solver_type == 2;
  const long n = nu + nv + (mono ? np : 0);
  const int bl = !per_x(m), br = !per_x(m), bb = !per_y(m), bt = !per_y(m);
  const int sync = per_x(m) || per_y(m);
  if (precond == 1) {
    for (long k = 0; k < n; ++k) dst[k] = fabs(B->diag[k]) > 1.0e-15 ? src[k] / B->diag[k] : src[k];
    return;
  }
  if (precond != 2 && precond != 4) {
    memcpy(dst, src, n * sizeof(double));
    return;
  }
  const long nh = (long)(nx + 2) * (ny + 2);
  double *xu = malloc(nh * sizeof(double)), *xv = malloc(nh * sizeof(double)), *xp = malloc(nh * sizeof(double));
  double *ru = malloc(nu * sizeof(double)), *rv = malloc(nv * sizeof(double));
  if (precond == 2 || !mono) {
    orc_ssor_block(m, u_nx, ny, B->uW, B->uE, B->uS, B->uN, B->uP, src, xu, 2, 1.0, sync);
    orc_ssor_block(m, nx, v_ny, B->vW, B->vE, B->vS, B->vN, B->vP, src + nu, xv, 2, 1.0, sync);
    if (mono) orc_ssor_block(m, nx, ny, B->sW, B->sE, B->sS, B->sN, B->sP, src + nu + nv, xp, 2, 1.0, sync);
  } else {
    orc_ssor_block(m, nx, ny, B->sW, B->sE, B->sS, B->sN, B->sP, src + nu + nv, xp, 12, 1.25, sync);
    double sum = 0.0;
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) sum = sum + xp[H(m, i, j)];
    const double mean = sum / (double)((long)nx * ny);
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) xp[H(m, i, j)] = xp[H(m, i, j)] - mean;
    if (sync) orc_exchange_halos_scalar(m, xp);
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= u_nx; ++i) {
        const long dof = (long)(j - 1) * u_nx + (i - 1);
        const int is_wall = (i == 1 && bl) || (i == u_nx && br);
        ru[dof] = is_wall ? src[dof] : src[dof] - (xp[H(m, i, j)] - xp[H(m, i - 1, j)]) * m->dy;
      }
    orc_ssor_block(m, u_nx, ny, B->uW, B->uE, B->uS, B->uN, B->uP, ru, xu, 4, 1.0, sync);
    for (int j = 1; j <= v_ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const long dof = (long)(j - 1) * nx + (i - 1);
        const int is_wall = (j == 1 && bb && wall(m->bc[2])) || (j == v_ny && bt && wall(m->bc[3])) ||
                          solid_v(m, i, j);
        rv[dof] = is_wall ? src[nu + dof] : src[nu + dof] - (xp[H(m, i, j)] - xp[H(m, i, j - 1)]) * m->dx;
      }
    orc_ssor_block(m, nx, v_ny, B->vW, B->vE, B->vS, B->vN, B->vP, rv, xv, 4, 1.0, sync);
  }
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= u_nx; ++i) dst[(long)(j - 1) * u_nx + (i - 1)] = xu[H(m, i, j)];
  for (int j = 1; j <= v_ny; ++j)
    for (int i = 1; i <= nx; ++i) dst[nu + (long)(j - 1) * nx + (i - 1)] = xv[H(m, i, j)];
  if (mono)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) dst[nu + nv + W(m, i, j)] = xp[H(m, i, j)];
  free(xu); free(xv); free(xp); free(ru); free(rv);
}

/* ns_assemble_rhs (src/module_navier_stokes.f90:785-1101), SP branch. */
void orc_mom_rhs(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                 double *rhs) {
  const int nx = m->nx, ny = m->ny;
  const double dx = m->dx, dy = m->dy, dvol = dx * dy;
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const int bl = !per_x(m), br = !per_x(m), bb = !per_y(m), bt = !per_y(m);
  const long nu = (long)u_nx * ny;
  const double at = gear(c) ? 1.5 : 1.0;
  const double *u = F->u, *v = F->v;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= u_nx; ++i) {
      const long dof = (long)(j - 1) * u_nx + (i - 1);
      const double rho_u = F->rho_u[H(m, i, j)];
      const double diag_bc = (at * rho_u / dt) * dvol;
      if (i == 1 && bl) {
        rhs[dof] = diag_bc * (m->bc[0] == BC_DIRICHLET ? c->bc_u[0] : 0.0);
        continue;
      }
      if (i == u_nx && br) {
        rhs[dof] = diag_bc * (m->bc[1] == BC_DIRICHLET ? c->bc_u[1] : 0.0);
        continue;
      }
      double val;
      if (gear(c)) {
        const double ui = rho_u * (2.0 * F->u_old[H(m, i, j)] - 0.5 * F->u_nm1[H(m, i, j)]);
        val = (ui / dt) * dvol;
      } else {
        val = (rho_u / dt) * F->u_old[H(m, i, j)] * dvol;
      }
      val = val + rho_u * c->gx * dvol;
      val = val + F->fx[H(m, i, j)] * dvol;
      if (c->conv_lin == 3) {
        const double du_dx = (u[H(m, i + 1, j)] - u[H(m, i - 1, j)]) / (2.0 * dx);
        const double du_dy = (u[H(m, i, j + 1)] - u[H(m, i, j - 1)]) / (2.0 * dy);
        const double vf = 0.25 * (v[H(m, i - 1, j)] + v[H(m, i, j)] + v[H(m, i - 1, j + 1)] + v[H(m, i, j + 1)]);
        val = val + rho_u * (u[H(m, i, j)] * du_dx + vf * du_dy) * dvol;
      }
      if (j == ny && bt && m->bc[3] == BC_DIRICHLET) {
        const double mu_N = F->mu_node[H(m, i, j + 1)];
        val = ny >= 2 ? val + mu_N * ((8.0 / 3.0) * c->bc_u[3] / dy) * dx
                      : val + mu_N * (2.0 * c->bc_u[3] / dy) * dx;
      }
      if (j == 1 && bb && m->bc[2] == BC_DIRICHLET) {
        const double mu_S = F->mu_node[H(m, i, j)];
        val = ny >= 2 ? val + mu_S * ((8.0 / 3.0) * c->bc_u[2] / dy) * dx
                      : val + mu_S * (2.0 * c->bc_u[2] / dy) * dx;
      }
      rhs[dof] = val;
    }
  for (int j = 1; j <= v_ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const long dof = nu + (long)(j - 1) * nx + (i - 1);
      const double rho_v = F->rho_v[H(m, i, j)];
      const double diag_bc = (at * rho_v / dt) * dvol;
      if (j == 1 && bb && wall(m->bc[2])) {
        rhs[dof] = diag_bc * (m->bc[2] == BC_DIRICHLET ? c->bc_v[2] : 0.0);
        continue;
      }
      if (j == v_ny && bt && wall(m->bc[3])) {
        rhs[dof] = diag_bc * (m->bc[3] == BC_DIRICHLET ? c->bc_v[3] : 0.0);
        continue;
      }
      double val;
      if (gear(c)) {
        const double vi = rho_v * (2.0 * F->v_old[H(m, i, j)] - 0.5 * F->v_nm1[H(m, i, j)]);
        val = (vi / dt) * dvol;
      } else {
        val = (rho_v / dt) * F->v_old[H(m, i, j)] * dvol;
      }
      val = val + rho_v * c->gy * dvol;
      val = val + F->fy[H(m, i, j)] * dvol;
      if (c->conv_lin == 3) {
        const double dv_dx = (v[H(m, i + 1, j)] - v[H(m, i - 1, j)]) / (2.0 * dx);
        const double dv_dy = (v[H(m, i, j + 1)] - v[H(m, i, j - 1)]) / (2.0 * dy);
        const double uf = 0.25 * (u[H(m, i, j - 1)] + u[H(m, i + 1, j - 1)] + u[H(m, i, j)] + u[H(m, i + 1, j)]);
        val = val + rho_v * (uf * dv_dx + v[H(m, i, j)] * dv_dy) * dvol;
      }
      if (i == 1 && bl && m->bc[0] == BC_DIRICHLET) {
        const double mu_W = F->mu_node[H(m, i, j)];
        val = val + mu_W * (2.0 * c->bc_v[0] / dx) * dy;
      }
      if (i == nx && br && m->bc[1] == BC_DIRICHLET) {
        const double mu_E = F->mu_node[H(m, i + 1, j)];
        val = val + mu_E * (2.0 * c->bc_v[1] / dx) * dy;
      }
      rhs[dof] = val;
    }
}

/* solve_cg (src/module_linear_solvers.f90:735-838) and solve_bicgstab2
 * (:361-579) on the momentum operator (SP) or the monolithic operator (tp:
 * its pressure work field); the preconditioner from the blocks of the
 * fields; returns the reference's iters. */
int orc_mom_solve2(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt, double *tu,
                   double *tv, double *tp, int lin_solver, int precond, double tol, int maxit, const double *b,
                   double *x, double *relres) {
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const long nu = (long)u_nx * m->ny, nv = (long)m->nx * v_ny, np = (long)m->nx * m->ny;
  const int mono = c->solver_type == 2;
  const long n = nu + nv + (mono ? np : 0);
  double *bk[16];
  for (int k = 0; k < 16; ++k) bk[k] = malloc((k == 0 ? n : (k <= 5 ? nu : (k <= 10 ? nv : np))) * sizeof(double));
  orc_mom_blocks_t B = {bk[0], bk[1], bk[2], bk[3], bk[4], bk[5], bk[6], bk[7], bk[8], bk[9], bk[10],
                        bk[11], bk[12], bk[13], bk[14], bk[15]};
  orc_mom_blocks(m, c, F, dt, &B);
#define MATVEC(src, dst)                                                      \
  do {                                                                        \
    if (mono) orc_mono_matvec(m, c, F, dt, tu, tv, tp, src, dst);             \
    else orc_mom_matvec(m, c, F, dt, tu, tv, src, dst);                       \
  } while (0)
#define PRECOND(src, dst) orc_mom_precond(m, c, precond, &B, src, dst)
  double *w[9];
  for (int k = 0; k < 9; ++k) w[k] = calloc(n, sizeof(double));
  int iters;
  double rr = 0.0;
  if (lin_solver == 3) {
    double *Ax = w[0], *r = w[1], *z = w[2], *p = w[3], *Ap = w[4];
    iters = maxit;
    MATVEC(x, Ax);
    for (long k = 0; k < n; ++k) r[k] = b[k] - Ax[k];
    const double r0 = sqrt(orc_dot(r, r, n));
    if (r0 < 1.0e-12) {
      iters = 0;
    } else {
      PRECOND(r, z);
      memcpy(p, z, n * sizeof(double));
      double rz_old = orc_dot(r, z, n);
      for (int it = 1; it <= maxit; ++it) {
        MATVEC(p, Ap);
        const double pAp = orc_dot(p, Ap, n);
        if (fabs(pAp) < 1.0e-30) break;
        const double alpha = rz_old / pAp;
        for (long k = 0; k < n; ++k) x[k] = x[k] + alpha * p[k];
        for (long k = 0; k < n; ++k) r[k] = r[k] - alpha * Ap[k];
        const double rn = sqrt(orc_dot(r, r, n));
        rr = rn / r0;
        if (rn / r0 < tol || rn < 1.0e-13) {
          iters = it;
          break;
        }
        PRECOND(r, z);
        const double rz_new = orc_dot(r, z, n);
        const double beta = rz_new / rz_old;
        rz_old = rz_new;
        for (long k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];
      }
    }
  } else {
    double *bh = w[0], *rt = w[1], *r0 = w[2], *r1 = w[3], *r2 = w[4], *u0 = w[5], *u1 = w[6], *u2 = w[7];
    double *tmp = w[8];
#define PA(src, dst)                                      \
  do {                                                    \
    MATVEC(src, tmp);                                     \
    PRECOND(tmp, dst);                                    \
  } while (0)
    iters = 0;
    PRECOND(b, bh);
    PA(x, r0);
    for (long k = 0; k < n; ++k) r0[k] = bh[k] - r0[k];
    const double r0n = sqrt(orc_dot(r0, r0, n));
    int iter = 0;
    if (r0n >= 1.0e-12) {
      memcpy(rt, r0, n * sizeof(double));
      double rho_0 = 1.0, rho_1, alpha = 0.0, omega = 1.0, beta, gamma;
      while (iter < maxit) {
        iter += 2;
        rho_0 = -omega * rho_0;
        rho_1 = orc_dot(rt, r0, n);
        if (fabs(rho_0) < 1.0e-300) break;
        beta = alpha * (rho_1 / rho_0);
        rho_0 = rho_1;
        for (long k = 0; k < n; ++k) u0[k] = r0[k] - beta * u0[k];
        PA(u0, u1);
        gamma = orc_dot(rt, u1, n);
        if (fabs(gamma) < 1.0e-300) break;
        alpha = rho_0 / gamma;
        for (long k = 0; k < n; ++k) r0[k] = r0[k] - alpha * u1[k];
        for (long k = 0; k < n; ++k) x[k] = x[k] + alpha * u0[k];
        PA(r0, r1);
        rho_1 = orc_dot(rt, r1, n);
        if (fabs(rho_0) < 1.0e-300) break;
        beta = alpha * (rho_1 / rho_0);
        rho_0 = rho_1;
        for (long k = 0; k < n; ++k) u0[k] = r0[k] - beta * u0[k];
        for (long k = 0; k < n; ++k) u1[k] = r1[k] - beta * u1[k];
        PA(u1, u2);
        gamma = orc_dot(rt, u2, n);
        if (fabs(gamma) < 1.0e-300) break;
        alpha = rho_0 / gamma;
        for (long k = 0; k < n; ++k) x[k] = x[k] + alpha * u0[k];
        for (long k = 0; k < n; ++k) r0[k] = r0[k] - alpha * u1[k];
        for (long k = 0; k < n; ++k) r1[k] = r1[k] - alpha * u2[k];
        PA(r1, r2);
        const double s11 = orc_dot(r1, r1, n), s12 = orc_dot(r2, r1, n), s22 = orc_dot(r2, r2, n);
        const double rhs1 = orc_dot(r0, r1, n), rhs2 = orc_dot(r0, r2, n);
        const double rn1 = sqrt(s11), rn2 = sqrt(s22);
        double g1, g2;
        if (rn1 < 1.0e-300) {
          break;
        } else if (rn2 < 1.0e-300) {
          g2 = 0.0;
          g1 = rhs1 / s11;
        } else {
          const double i1 = 1.0 / rn1, i2 = 1.0 / rn2;
          const double c12 = s12 * i1 * i2, b1 = rhs1 * i1, b2 = rhs2 * i2;
          const double det = 1.0 - c12 * c12;
          if (fabs(det) < 1.0e-14) {
            g2 = 0.0;
            g1 = b1 * i1;
          } else {
            g1 = (b1 - c12 * b2) / (det * rn1);
            g2 = (b2 - c12 * b1) / (det * rn2);
          }
        }
        for (long k = 0; k < n; ++k) x[k] = x[k] + g1 * r0[k] + g2 * r1[k];
        for (long k = 0; k < n; ++k) r0[k] = r0[k] - (g1 * r1[k] + g2 * r2[k]);
        for (long k = 0; k < n; ++k) u0[k] = u0[k] - (g1 * u1[k] + g2 * u2[k]);
        if (g2 != 0.0) omega = g2;
        else if (fabs(g1) > 1.0e-300) omega = g1;
        else omega = 1.0;
        const double rn = sqrt(orc_dot(r0, r0, n));
        rr = rn / r0n;
        if (rn != rn || rn > 1.0e30) break;
        if (rn / r0n < tol || rn < 1.0e-12) break;
      }
    }
    iters = iter;
#undef PA
  }
#undef MATVEC
#undef PRECOND
  if (relres) *relres = rr;
  for (int k = 0; k < 9; ++k) free(w[k]);
  for (int k = 0; k < 16; ++k) free(bk[k]);
  return iters;
}

int orc_mom_solve(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                  const double *diag, double *tu, double *tv, int lin_solver, int precond, double tol,
                  int maxit, const double *b, double *x, double *relres) {
  (void)diag;
  return orc_mom_solve2(m, c, F, dt, tu, tv, NULL, lin_solver, precond, tol, maxit, b, x, relres);
}

/* The momentum block of sim_solve_standard_projection
 * (src/module_simulation.f90:2454-2469): diagonal, right-hand side, pack
 * (initial guess = the current u, v), solve, unpack_vector(.false.), then
 * sync_u, sync_v and apply_boundary_conditions(.false.) once more.  u, v are
 * updated in place (the reference's fields%u / fields%v, which are also the
 * fields_old of the operator during the solve: F->u, F->v must be copies). */
int orc_sp_predictor(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt,
                     double *tu, double *tv, int lin_solver, int precond, double tol, int maxit,
                     double *u, double *v, double *relres) {
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const long n = (long)u_nx * m->ny + (long)m->nx * v_ny;
  double *rhs = malloc(n * sizeof(double));
  double *x = malloc(n * sizeof(double));
  orc_mom_rhs(m, c, F, dt, rhs);
  orc_mom_pack(m, u, v, x);
  const int it = orc_mom_solve2(m, c, F, dt, tu, tv, NULL, lin_solver, precond, tol, maxit, rhs, x, relres);
  orc_mom_unpack(m, c, x, u, v, 0);
  orc_exchange_halos_scalar(m, u);
  orc_exchange_halos_scalar(m, v);
  orc_velocity_bc(m, c, u, v, 0);
  free(rhs);
  free(x);
  return it;
}

/* The monolithic branch of sim_run's solve (src/module_simulation.f90:
 * 553-583 with uzawa_iter 1): assemble_preconditioner_blocks, assemble_rhs
 * (zero continuity rows), pack (u, v, p), solve_linear_system,
 * unpack_vector(.false.) into u, v, p.  F->u, F->v are the fields_old. */
int orc_mono_step(const orc_mesh *m, const orc_mom_cfg *c, const orc_mom_fields *F, double dt, double *tu,
                  double *tv, double *tp, int lin_solver, int precond, double tol, int maxit, double *u, double *v,
                  double *p, double *relres) {
  int u_nx, v_ny;
  orc_mom_sizes(m, &u_nx, &v_ny);
  const long nuv = (long)u_nx * m->ny + (long)m->nx * v_ny, n = nuv + (long)m->nx * m->ny;
  double *rhs = calloc(n, sizeof(double)), *x = malloc(n * sizeof(double));
  orc_mom_rhs(m, c, F, dt, rhs);
  orc_mono_pack(m, u, v, p, x);
  const int it = orc_mom_solve2(m, c, F, dt, tu, tv, tp, lin_solver, precond, tol, maxit, rhs, x, relres);
  orc_mono_unpack(m, c, x, u, v, p, 0);
  free(rhs);
  free(x);
  return it;
}

/* ------------------------------------------------------------ output */

/* vtk_write_diagnostics (src/module_vtk_writer.f90:399-531), one rank, serial
 * sums: out[8] = center_of_mass_y, rise_velocity, bubble_volume,
 * kinetic_energy, max_u, max_div, avg_rho_gas, avg_p_gas.  T may be NULL
 * (300 everywhere). */
void orc_history(const orc_mesh *m, const double *C, const double *u, const double *v, const double *p,
                 const double *rho, const double *T, double *out) {
  double vol = 0.0, gvol = 0.0, yint = 0.0, vint = 0.0, mgas = 0.0, pgas = 0.0, ke = 0.0, mu = 0.0, md = 0.0;
  const double dvol = m->dx * m->dy;
  for (int j = 1; j <= m->ny; ++j) {
    const double y_pos = ((double)(j - 1) + 0.5) * m->dy;
    for (int i = 1; i <= m->nx; ++i) {
      const double c_val = C[H(m, i, j)];
      const double ind_bubble = 1.0 - c_val;
      vol = vol + ind_bubble * dvol;
      yint = yint + y_pos * ind_bubble * dvol;
      const double ind_gas = 1.0 - c_val;
      const double Tc = T ? T[H(m, i, j)] : 300.0;
      const double rho_gas_c = p[H(m, i, j)] / (287.05 * dmax(1.0, Tc));
      mgas = mgas + ind_gas * rho_gas_c * dvol;
      pgas = pgas + ind_gas * p[H(m, i, j)] * dvol;
      gvol = gvol + ind_gas * dvol;
      const double uc = 0.5 * (u[H(m, i, j)] + u[H(m, i + 1, j)]);
      const double vc = 0.5 * (v[H(m, i, j)] + v[H(m, i, j + 1)]);
      vint = vint + vc * ind_bubble * dvol;
      const double speed2 = uc * uc + vc * vc;
      const double u2p = 0.5 * (u[H(m, i, j)] * u[H(m, i, j)] + u[H(m, i + 1, j)] * u[H(m, i + 1, j)]);
      const double v2p = 0.5 * (v[H(m, i, j)] * v[H(m, i, j)] + v[H(m, i, j + 1)] * v[H(m, i, j + 1)]);
      ke = ke + 0.5 * rho[H(m, i, j)] * (u2p + v2p) * dvol;
      const double speed = sqrt(speed2);
      if (speed > mu) mu = speed;
      const double dv = fabs((u[H(m, i + 1, j)] - u[H(m, i, j)]) / m->dx + (v[H(m, i, j + 1)] - v[H(m, i, j)]) / m->dy);
      if (dv > md) md = dv;
    }
  }
  out[0] = vol > 1.0e-12 ? yint / vol : 0.0;
  out[1] = vol > 1.0e-12 ? vint / vol : 0.0;
  out[2] = vol;
  out[3] = ke;
  out[4] = mu;
  out[5] = md;
  out[6] = gvol > 1.0e-12 ? mgas / gvol : 0.0;
  out[7] = gvol > 1.0e-12 ? pgas / gvol : 0.0;
}
