// Adjoint of the variable-density pressure solve (no reference counterpart).
//
// Forward: A(rho) x = b with the face coefficients of the SP assembly
// (src/module_simulation.f90:2483-2533), k_f = dt*dy / (rho_f*dx) on x faces
// (dt*dx / (rho_f*dy) on y faces), rho_f = (rho_L + rho_R)/2
// (src/module_interface_vof.f90:574-590).  For an objective J(x):
//   A^T lam = dJ/dx                      (A is symmetric: the same solver)
//   dJ/db = lam
//   dJ/dk_f = -(lam_R - lam_L)(x_R - x_L)  for an interior face L|R,
//   dJ/dk_f = -lam_c x_c                   for a non-wall, non-periodic edge face,
//   dJ/d(log rho_c) = rho_c * sum_f dJ/dk_f * (-k_f/rho_f) * 1/2
// over the faces f whose mean uses rho_c.  Halo densities are the copies the
// reference's halo/BC fill makes (periodic wrap, zero-gradient otherwise), so
// a boundary face's halo half is credited to the interior cell it copies.
#include "covof_internal.h"

namespace covof {

int poisson_solve(Ctx &c, const double *b, double *x, int lin_solver, int precond, double tol,
                  int maxit, int precision, int *iters, double *relres);

namespace {

// the kx layout of poisson.hip: rows of nx+1 faces padded to a multiple of 64
__device__ __forceinline__ long kxi(int nx, int i, int j) {
  return (long)(j - 1) * (((long)(nx + 1) + 63) & ~63L) + (i - 1);
}
__device__ __forceinline__ long kyi(int nx, int i, int j) { return (long)(j - 1) * nx + (i - 1); }

// dJ/dk * dk/drho_f * 1/2 for every face (one value per face, kx / ky layout).
__global__ void k_face_sens(int nx, int ny, int perW, int perE, int perS, int perN,
                            const double *__restrict__ kx, const double *__restrict__ ky,
                            const double *__restrict__ rho, const double *__restrict__ x,
                            const double *__restrict__ lam, double *__restrict__ gx,
                            double *__restrict__ gy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;  // 1..nx+1
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;  // 1..ny+1
  auto W = [&](int a, int c) { return (long)(c - 1) * nx + (a - 1); };
  auto Hh = [&](int a, int c) { return (long)c * (nx + 2) + a; };
  auto dk_drho = [&](double k, double rl, double rr) {
    const double rf = 0.5 * (rl + rr);
    return (rf < 1.0e-12) ? 0.0 : -k / rf * 0.5;
  };
  if (i <= nx + 1 && j <= ny) {
    const double k = kx[kxi(nx, i, j)];
    double g = 0.0;
    if (k != 0.0) {
      double dJdk;
      if (i > 1 && i <= nx) {
        dJdk = -(lam[W(i, j)] - lam[W(i - 1, j)]) * (x[W(i, j)] - x[W(i - 1, j)]);
      } else if (i == 1) {
        const double xn = perW ? x[W(nx, j)] : 0.0;
        dJdk = -lam[W(1, j)] * (x[W(1, j)] - xn);
      } else {
        const double xn = perE ? x[W(1, j)] : 0.0;
        dJdk = -lam[W(nx, j)] * (x[W(nx, j)] - xn);
      }
      g = dJdk * dk_drho(k, rho[Hh(i - 1, j)], rho[Hh(i, j)]);
    }
    gx[kxi(nx, i, j)] = g;
  }
  if (i <= nx && j <= ny + 1) {
    const double k = ky[kyi(nx, i, j)];
    double g = 0.0;
    if (k != 0.0) {
      double dJdk;
      if (j > 1 && j <= ny) {
        dJdk = -(lam[W(i, j)] - lam[W(i, j - 1)]) * (x[W(i, j)] - x[W(i, j - 1)]);
      } else if (j == 1) {
        const double xn = perS ? x[W(i, ny)] : 0.0;
        dJdk = -lam[W(i, 1)] * (x[W(i, 1)] - xn);
      } else {
        const double xn = perN ? x[W(i, 1)] : 0.0;
        dJdk = -lam[W(i, ny)] * (x[W(i, ny)] - xn);
      }
      g = dJdk * dk_drho(k, rho[Hh(i, j - 1)], rho[Hh(i, j)]);
    }
    gy[kyi(nx, i, j)] = g;
  }
}

// Gather the face sensitivities to cells: cell (i,j) takes its four faces,
// plus the halo half of a boundary face it is the copy source of.
__global__ void k_cell_grad(int nx, int ny, int perx, int pery, const double *__restrict__ gx,
                            const double *__restrict__ gy, const double *__restrict__ rho,
                            double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > nx || j > ny) return;
  double s = gx[kxi(nx, i, j)] + gx[kxi(nx, i + 1, j)] + gy[kyi(nx, i, j)] + gy[kyi(nx, i, j + 1)];
  // halo copies: rho(0,j) = rho(nx,j) (periodic) or rho(1,j) (zero gradient), etc.
  if (perx) {
    if (i == nx) s += gx[kxi(nx, 1, j)];
    if (i == 1) s += gx[kxi(nx, nx + 1, j)];
  } else {
    if (i == 1) s += gx[kxi(nx, 1, j)];
    if (i == nx) s += gx[kxi(nx, nx + 1, j)];
  }
  if (pery) {
    if (j == ny) s += gy[kyi(nx, i, 1)];
    if (j == 1) s += gy[kyi(nx, i, ny + 1)];
  } else {
    if (j == 1) s += gy[kyi(nx, i, 1)];
    if (j == ny) s += gy[kyi(nx, i, ny + 1)];
  }
  out[(long)(j - 1) * nx + (i - 1)] = s * rho[(long)j * (nx + 2) + i];
}

__global__ void k_zero64(long n, double *x) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) x[t] = 0.0;
}

}  // namespace

int poisson_adjoint(Ctx &c, const double *rho, double dt, const double *x, const double *w,
                    double *dJ_db, double *dJ_dlogrho, int lin_solver, int precond, double tol,
                    int maxit, int precision, int *iters, double *relres) {
  (void)dt;
  if (!c.ps.assembled) return COVOF_BAD_ARGUMENT;
  const Mesh &m = c.m;
  const long n = (long)m.nx * m.ny;
  // lam = A^-1 w from lam = 0 (w must not alias the solver's work vectors)
  hipLaunchKernelGGL(k_zero64, dim3(1024), dim3(256), 0, c.stream, n, dJ_db);
  COVOF_CHECK_LAUNCH();
  const int st = poisson_solve(c, w, dJ_db, lin_solver, precond, tol, maxit, precision, iters, relres);
  if (st == COVOF_BAD_ARGUMENT || st == COVOF_HIP_ERROR) return st;
  const dim3 blk(64, 4), grd((m.nx + 1 + 63) / 64, (m.ny + 1 + 3) / 4);
  hipLaunchKernelGGL(k_face_sens, grd, blk, 0, c.stream, m.nx, m.ny, m.bc[0] == BC_PERIODIC,
                     m.bc[1] == BC_PERIODIC, m.bc[2] == BC_PERIODIC, m.bc[3] == BC_PERIODIC,
                     c.ps.kx, c.ps.ky, rho, x, dJ_db, c.ps.fx, c.ps.fy);
  COVOF_CHECK_LAUNCH();
  const dim3 grd2((m.nx + 63) / 64, (m.ny + 3) / 4);
  hipLaunchKernelGGL(k_cell_grad, grd2, blk, 0, c.stream, m.nx, m.ny, m.perx, m.pery, c.ps.fx, c.ps.fy,
                     rho, dJ_dlogrho);
  COVOF_CHECK_LAUNCH();
  return st;
}

}  // namespace covof
