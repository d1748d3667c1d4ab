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

// ---------------------------------------------- adjoint of the projection
//
// Forward (src/module_simulation.f90:2477-2654): b = -div(u*) (:2485-2487),
// A x = b, p = x - mean(x) (:2577-2585), and the correction of every
// non-wall face, u = u* - (dt/rho_u) (p_e - p_w)/dx (:2590-2654; v alike),
// whose pressure indices wrap on periodic sides and repeat the edge cell
// otherwise (a zero gradient).  For J(u, v, p) with w_u = dJ/du, w_v = dJ/dv,
// w_p = dJ/dp:
//   g = w_p + sum over corrected faces of the gradient's adjoint (per cell)
//   A lam = g - mean(g)             (the mean removal is a symmetric projector)
//   dJ/du* = w_u + dy (lam_E - lam_W) ... the divergence's transpose,
//   dJ/drho_f = w_f dt grad_f / rho_f^2 (the correction) + the operator's
//   face sensitivity, each credited half to the two cells of the face mean.

// The correction's reading of face f of row j (k_correct's rule): whether it
// is corrected, its west / east pressure cells, dt/(rho_f d).
struct FaceCorr {
  bool on;
  int cw, ce;
  double c;
};
__device__ __forceinline__ FaceCorr u_corr(int nx, int perx, int wallL, int wallR, const double *rho_u, int f,
                                           int j, double dt, double dx) {
  FaceCorr r{false, 0, 0, 0.0};
  const int u_nx = perx ? nx : nx + 1;
  if (f < 1 || f > u_nx) return r;
  if ((f == 1 && !perx && wallL) || (f == u_nx && !perx && wallR)) return r;
  double rf = rho_u[(long)j * (nx + 2) + f];
  if (rf < 1.0e-12) rf = 1.0;
  r.on = true;
  r.cw = f > 1 ? f - 1 : (perx ? nx : 1);
  r.ce = f <= nx ? f : (perx ? 1 : nx);
  r.c = dt / rf / dx;
  return r;
}

// g(i,j) = w_p(i,j) + the correction's pressure adjoint (cells of the work
// layout), per-block partial sums of g in part[blockIdx].
__global__ void __launch_bounds__(256) k_proj_adj_g(int nx, int ny, int perx, int pery, int wallL, int wallR,
                                                    int wallB, int wallT, double dx, double dy, double dt,
                                                    const double *__restrict__ rho_u,
                                                    const double *__restrict__ rho_v,
                                                    const double *__restrict__ w_u, const double *__restrict__ w_v,
                                                    const double *__restrict__ w_p, double *__restrict__ g,
                                                    double *__restrict__ part) {
  __shared__ double sh[4];
  const long n = (long)nx * ny;
  double acc = 0.0;
  const int u_nx = perx ? nx : nx + 1, v_ny = pery ? ny : ny + 1;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int j = (int)(t / nx) + 1, i = (int)(t - (long)(j - 1) * nx) + 1;
    double s = 0.0;
    if (w_p) {
      // dJ/dp of the cell and of its periodic halo copies (the wrap copies
      // x first, rows 0..ny+1, then y, columns 0..nx+1)
      auto Wp = [&](int a, int b) { return w_p[(long)b * (nx + 2) + a]; };
      auto col = [&](int a) {  // p(a, j) and the y-wrap copies of it
        double t = Wp(a, j);
        if (pery && j == ny) t += Wp(a, 0);
        if (pery && j == 1) t += Wp(a, ny + 1);
        return t;
      };
      s = col(i);
      if (perx && i == nx) s += col(0);
      if (perx && i == 1) s += col(nx + 1);
    }
    // x faces that can read p(i, j): i, i + 1, the periodic wrap face 1 (cell
    // nx) and the last corrected face (cell 1); each counted once
    int fl[4] = {i, i + 1, i == nx ? 1 : -1, i == 1 ? u_nx : -1};
    for (int q = 0; q < 4; ++q) {
      const int f = fl[q];
      bool dup = f < 0;
      for (int r = 0; r < q; ++r) dup = dup || fl[r] == f;
      if (dup) continue;
      const FaceCorr fc = u_corr(nx, perx, wallL, wallR, rho_u, f, j, dt, dx);
      if (!fc.on || fc.cw == fc.ce) continue;  // an edge face reads one cell twice: no gradient
      const double wv = w_u[(long)j * (nx + 2) + f];
      if (fc.cw == i) s += fc.c * wv;
      if (fc.ce == i) s -= fc.c * wv;
    }
    int fr[4] = {j, j + 1, j == ny ? 1 : -1, j == 1 ? v_ny : -1};
    for (int q = 0; q < 4; ++q) {
      const int f = fr[q];
      bool dup = f < 0;
      for (int r = 0; r < q; ++r) dup = dup || fr[r] == f;
      if (dup) continue;
      if (f < 1 || f > v_ny) continue;
      if ((f == 1 && !pery && wallB) || (f == v_ny && !pery && wallT)) continue;
      double rf = rho_v[(long)f * (nx + 2) + i];
      if (rf < 1.0e-12) rf = 1.0;
      const int cs = f > 1 ? f - 1 : (pery ? ny : 1);
      const int cn = f <= ny ? f : (pery ? 1 : ny);
      if (cs == cn) continue;
      const double c = dt / rf / dy;
      const double wv = w_v[(long)f * (nx + 2) + i];
      if (cs == j) s += c * wv;
      if (cn == j) s -= c * wv;
    }
    g[t] = s;
    acc += s;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// g -= mean(g) (every block sums the nblk partials in the same order)
__global__ void __launch_bounds__(256) k_proj_adj_center(long n, int nblk, const double *__restrict__ part,
                                                         double *__restrict__ g) {
  __shared__ double tot;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[b];
    tot = s / (double)n;
  }
  __syncthreads();
  const double mean = tot;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) g[t] = g[t] - mean;
}

// dJ/du*, dJ/dv* (halo layout): the direct term and the divergence's
// transpose; the correction's face-density terms added to the operator's
// face sensitivities (kx / ky layout, half per cell of the face mean).
__global__ void k_proj_adj_vel(int nx, int ny, int perx, int pery, int wallL, int wallR, int wallB, int wallT,
                               double dx, double dy, double dt, const double *__restrict__ rho_u,
                               const double *__restrict__ rho_v, const double *__restrict__ x,
                               const double *__restrict__ lam, const double *__restrict__ w_u,
                               const double *__restrict__ w_v, double *__restrict__ du, double *__restrict__ dv,
                               double *__restrict__ gx, double *__restrict__ gy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nx+1
  const int j = blockIdx.y * blockDim.y + threadIdx.y;  // 0..ny+1
  if (i > nx + 1 || j > ny + 1) return;
  const long h = (long)j * (nx + 2) + i;
  auto W = [&](int a, int c) { return (long)(c - 1) * nx + (a - 1); };
  double su = w_u[h], sv = w_v[h];
  if (j >= 1 && j <= ny && i >= 1) {
    if (i <= nx) su += dy * lam[W(i, j)];
    if (i >= 2) su -= dy * lam[W(i - 1, j)];
    // the correction's dependence on rho_u of this face
    const FaceCorr fc = u_corr(nx, perx, wallL, wallR, rho_u, i, j, dt, dx);
    if (fc.on && i <= nx + 1) {
      const double rf = rho_u[h];
      if (rf >= 1.0e-12) {
        const double grad = (x[W(fc.ce, j)] - x[W(fc.cw, j)]) / dx;
        gx[kxi(nx, i, j)] += 0.5 * (w_u[h] * (dt * grad) / (rf * rf));
      }
    }
  }
  if (i >= 1 && i <= nx && j >= 1) {
    if (j <= ny) sv += dx * lam[W(i, j)];
    if (j >= 2 && j - 1 <= ny) sv -= dx * lam[W(i, j - 1)];
    const int v_ny = pery ? ny : ny + 1;
    const bool on = j <= v_ny && !((j == 1 && !pery && wallB) || (j == v_ny && !pery && wallT));
    if (on) {
      const double rf = rho_v[h];
      if (rf >= 1.0e-12) {
        const int cs = j > 1 ? j - 1 : (pery ? ny : 1);
        const int cn = j <= ny ? j : (pery ? 1 : ny);
        const double grad = (x[W(i, cn)] - x[W(i, cs)]) / dy;
        gy[kyi(nx, i, j)] += 0.5 * (w_v[h] * (dt * grad) / (rf * rf));
      }
    }
  }
  du[h] = su;
  dv[h] = sv;
}

// Scale of the transpose solve's right-hand side: the Krylov loop's
// breakdown guards are absolute (|r.z|, |d.Ad| >= 1e-35), so a right-hand
// side near rounding level (an objective whose velocity gradient the
// correction nearly annihilates, as for the kinetic energy) is solved at
// unit norm and the solution scaled back.  Per-block partials of g^2:
__global__ void __launch_bounds__(256) k_sumsq_part(long n, const double *__restrict__ g, double *__restrict__ part) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) acc += g[t] * g[t];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// g /= |g| (every block sums the nblk partials in the same order); the norm
// (1 when it is zero or not finite, g then unchanged) to *s by block 0.
__global__ void __launch_bounds__(256) k_scale_unit(long n, int nblk, const double *__restrict__ part,
                                                    double *__restrict__ g, double *__restrict__ s) {
  __shared__ double nrm;
  if (threadIdx.x == 0) {
    double q = 0.0;
    for (int b = 0; b < nblk; ++b) q += part[b];
    const double r = sqrt(q);
    nrm = (r > 0.0 && r < 1.0e300) ? r : 1.0;
    if (blockIdx.x == 0) *s = nrm;
  }
  __syncthreads();
  const double inv = 1.0 / nrm;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) g[t] = g[t] * inv;
}

__global__ void k_scale_by(long n, const double *__restrict__ s, double *__restrict__ x) {
  const double f = *s;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) x[t] = x[t] * f;
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

int projection_adjoint(Ctx &c, const double *rho, const double *rho_u, const double *rho_v, double dt,
                       const double *x, const double *w_u, const double *w_v, const double *w_p, double *dJ_du,
                       double *dJ_dv, double *dJ_dlogrho, int precond, double tol, int maxit, int precision,
                       int *iters, double *relres) {
  if (!c.ps.assembled || c.obs.enabled) return COVOF_BAD_ARGUMENT;  // no obstacle faces in the adjoint
  const Mesh &m = c.m;
  const long n = (long)m.nx * m.ny;
  Poisson &P = c.ps;
  const int wl = m.bc[0] == BC_DIRICHLET || m.bc[0] == BC_FREESLIP;
  const int wr = m.bc[1] == BC_DIRICHLET || m.bc[1] == BC_FREESLIP;
  const int wb = m.bc[2] == BC_DIRICHLET || m.bc[2] == BC_FREESLIP;
  const int wt = m.bc[3] == BC_DIRICHLET || m.bc[3] == BC_FREESLIP;
  // g -> t4 (mean-free), lam -> t5 (the CG solves leave both alone); 256
  // partial sums past the first two partial buffers (the allocation holds
  // max(12 nblocks, 4608) values)
  double *g = P.t4, *lam = P.t5, *part = P.part + 2L * P.nblocks * 4;
  const int nblk = 256;
  hipLaunchKernelGGL(k_proj_adj_g, dim3(nblk), dim3(256), 0, c.stream, m.nx, m.ny, m.perx, m.pery, wl, wr, wb, wt,
                     m.dx, m.dy, dt, rho_u, rho_v, w_u, w_v, w_p, g, part);
  hipLaunchKernelGGL(k_proj_adj_center, dim3(1024), dim3(256), 0, c.stream, n, nblk, (const double *)part, g);
  // the solve at unit right-hand side norm (slot 57 keeps the norm)
  hipLaunchKernelGGL(k_sumsq_part, dim3(nblk), dim3(256), 0, c.stream, n, (const double *)g, part);
  hipLaunchKernelGGL(k_scale_unit, dim3(nblk), dim3(256), 0, c.stream, n, nblk, (const double *)part, g,
                     P.scal + 57);
  hipLaunchKernelGGL(k_zero64, dim3(1024), dim3(256), 0, c.stream, n, lam);
  COVOF_CHECK_LAUNCH();
  const int st = poisson_solve(c, g, lam, COVOF_LINEAR_CG, precond, tol, maxit, precision, iters, relres);
  if (st == COVOF_BAD_ARGUMENT || st == COVOF_HIP_ERROR) return st;
  hipLaunchKernelGGL(k_scale_by, dim3(1024), dim3(256), 0, c.stream, n, (const double *)(P.scal + 57), lam);
  const dim3 blk(64, 4), grd((m.nx + 1 + 63) / 64, (m.ny + 1 + 3) / 4);
  // the operator's face sensitivities (lam, x), then the correction's terms
  hipLaunchKernelGGL(k_face_sens, grd, blk, 0, c.stream, m.nx, m.ny, m.bc[0] == BC_PERIODIC,
                     m.bc[1] == BC_PERIODIC, m.bc[2] == BC_PERIODIC, m.bc[3] == BC_PERIODIC,
                     P.kx, P.ky, rho, x, (const double *)lam, P.fx, P.fy);
  const dim3 grdh((m.nx + 2 + 63) / 64, (m.ny + 2 + 3) / 4);
  hipLaunchKernelGGL(k_proj_adj_vel, grdh, blk, 0, c.stream, m.nx, m.ny, m.perx, m.pery, wl, wr, wb, wt, m.dx, m.dy,
                     dt, rho_u, rho_v, x, (const double *)lam, w_u, w_v, dJ_du, dJ_dv, P.fx, P.fy);
  const dim3 grd2((m.nx + 63) / 64, (m.ny + 3) / 4);
  hipLaunchKernelGGL(k_cell_grad, grd2, blk, 0, c.stream, m.nx, m.ny, m.perx, m.pery, P.fx, P.fy, rho,
                     dJ_dlogrho);
  COVOF_CHECK_LAUNCH();
  return st;
}

// ------------------------------------------- pieces of the coupled step adjoint
//
// The multi-step adjoint of the kinetic energy chains, per step, the
// adjoint of the transport (vof.hip), of the projection (above), of the
// density update and of the inertial predictor; these are the last two and
// the objective.

namespace {

// rho of update_physical_properties (src/module_interface_vof.f90:552-553)
// from the clamped C, and whether the clamp passed C through (derivative 1).
__device__ __forceinline__ double rho_of(double c, double rho1, double rho2) {
  const double t = 1.0 < c ? 1.0 : c;
  const double cc = 0.0 > t ? 0.0 : t;
  return cc * rho1 + (1.0 - cc) * rho2;
}
__device__ __forceinline__ bool clamp_passes(double c) { return !(c > 1.0) && !(c < 0.0); }

// Per-block partials of the kinetic energy of the history row
// (src/module_vtk_writer.f90:399-531): 0.5 rho (0.5 (uw^2 + ue^2) +
// 0.5 (vs^2 + vn^2)) dx dy summed over the cells.
__global__ void __launch_bounds__(256) k_ke_part(int nx, int ny, double dvol, const double *__restrict__ C,
                                                 const double *__restrict__ u, const double *__restrict__ v,
                                                 double rho1, double rho2, double *__restrict__ part) {
  __shared__ double sh[4];
  const long n = (long)nx * ny;
  double acc = 0.0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int j = (int)(t / nx) + 1, i = (int)(t - (long)(j - 1) * nx) + 1;
    const long h = (long)j * (nx + 2) + i;
    const double uw = u[h], ue = u[h + 1], vs = v[h], vn = v[h + nx + 2];
    const double u2p = 0.5 * (uw * uw + ue * ue), v2p = 0.5 * (vs * vs + vn * vn);
    acc += 0.5 * rho_of(C[h], rho1, rho2) * (u2p + v2p) * dvol;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

__global__ void k_ke_total(int nblk, const double *__restrict__ part, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[b];
  *out = s;
}

// dKE/dC (cells), dKE/du, dKE/dv (faces), halo layout; null outputs skipped.
__global__ void k_ke_grad(int nx, int ny, double dvol, const double *__restrict__ C, const double *__restrict__ u,
                          const double *__restrict__ v, double rho1, double rho2, double *__restrict__ dC,
                          double *__restrict__ du, double *__restrict__ dv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx + 1 || j > ny + 1) return;
  const long h = (long)j * (nx + 2) + i;
  const bool cell = i >= 1 && i <= nx && j >= 1 && j <= ny;
  if (dC) {
    double g = 0.0;
    if (cell && clamp_passes(C[h])) {
      const double uw = u[h], ue = u[h + 1], vs = v[h], vn = v[h + nx + 2];
      g = 0.5 * (0.5 * (uw * uw + ue * ue) + 0.5 * (vs * vs + vn * vn)) * dvol * (rho1 - rho2);
    }
    dC[h] = g;
  }
  if (du) {
    double g = 0.0;
    if (j >= 1 && j <= ny && i >= 1) {
      double r = 0.0;
      if (i >= 2) r += rho_of(C[h - 1], rho1, rho2);
      if (i <= nx) r += rho_of(C[h], rho1, rho2);
      g = 0.5 * u[h] * dvol * r;
    }
    du[h] = g;
  }
  if (dv) {
    double g = 0.0;
    if (i >= 1 && i <= nx && j >= 1) {
      double r = 0.0;
      if (j >= 2) r += rho_of(C[h - (nx + 2)], rho1, rho2);
      if (j <= ny) r += rho_of(C[h], rho1, rho2);
      g = 0.5 * v[h] * dvol * r;
    }
    dv[h] = g;
  }
}

// dJ/dC += dJ/d(log rho) (rho1 - rho2) / rho on the interior cells.
__global__ void k_props_adj(int nx, int ny, const double *__restrict__ C, double rho1, double rho2,
                            const double *__restrict__ g, double *__restrict__ dC) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > nx || j > ny) return;
  const long h = (long)j * (nx + 2) + i;
  const double c = C[h];
  if (!clamp_passes(c)) return;
  dC[h] += g[(long)(j - 1) * nx + (i - 1)] * (rho1 - rho2) / rho_of(c, rho1, rho2);
}

// The face sets of the predictor: x faces (1..nx+1, 1..ny), y faces
// (1..nx, 1..ny+1), without the wall faces (zero normal velocity).
struct PredFaces {
  int nx, ny, wl, wr, wb, wt;
  __device__ bool xface(int i, int j) const {
    return j >= 1 && j <= ny && i >= 1 && i <= nx + 1 && !(i == 1 && wl) && !(i == nx + 1 && wr);
  }
  __device__ bool yface(int i, int j) const {
    return i >= 1 && i <= nx && j >= 1 && j <= ny + 1 && !(j == 1 && wb) && !(j == ny + 1 && wt);
  }
};

// u* of the momentum block with convection and viscosity off: the BDF2
// inertia (4 u^n - u^{n-1} + 2 dt g) / 3, or explicit Euler u^n + dt g on
// the first step; zero on wall faces and outside the face sets.
__global__ void k_pred(PredFaces F, int first, double dt, double gx, double gy, const double *__restrict__ u,
                       const double *__restrict__ v, const double *__restrict__ un1, const double *__restrict__ vn1,
                       double *__restrict__ us, double *__restrict__ vs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > F.nx + 1 || j > F.ny + 1) return;
  const long h = (long)j * (F.nx + 2) + i;
  double a = 0.0, b = 0.0;
  if (F.xface(i, j)) a = first ? u[h] + dt * gx : (4.0 * u[h] - un1[h] + 2.0 * dt * gx) / 3.0;
  if (F.yface(i, j)) b = first ? v[h] + dt * gy : (4.0 * v[h] - vn1[h] + 2.0 * dt * gy) / 3.0;
  us[h] = a;
  vs[h] = b;
}

// The predictor's transpose: du += a w_u, dun1 += b w_u on its faces.
__global__ void k_pred_adj(PredFaces F, int first, const double *__restrict__ wu, const double *__restrict__ wv,
                           double *__restrict__ du, double *__restrict__ dv, double *__restrict__ dun1,
                           double *__restrict__ dvn1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > F.nx + 1 || j > F.ny + 1) return;
  const long h = (long)j * (F.nx + 2) + i;
  const double a = first ? 1.0 : 4.0 / 3.0, b = -1.0 / 3.0;
  if (F.xface(i, j)) {
    du[h] += a * wu[h];
    if (!first) dun1[h] += b * wu[h];
  }
  if (F.yface(i, j)) {
    dv[h] += a * wv[h];
    if (!first) dvn1[h] += b * wv[h];
  }
}

PredFaces pred_faces(const Mesh &m) {
  auto wall = [](int b) { return b == BC_DIRICHLET || b == BC_FREESLIP ? 1 : 0; };
  return PredFaces{m.nx, m.ny, wall(m.bc[0]), wall(m.bc[1]), wall(m.bc[2]), wall(m.bc[3])};
}

}  // namespace

// Sides must be walls (DIRICHLET / FREESLIP) or periodic pairs.
static bool closed_or_periodic(const Mesh &m) {
  for (int s = 0; s < 4; ++s)
    if (m.bc[s] != BC_DIRICHLET && m.bc[s] != BC_FREESLIP && m.bc[s] != BC_PERIODIC) return false;
  return true;
}

int kinetic_energy(Ctx &c, const double *C, const double *u, const double *v, double rho1, double rho2, double *ke,
                   double *dC, double *du, double *dv) {
  const Mesh &m = c.m;
  Poisson &P = c.ps;
  const double dvol = m.dx * m.dy;
  const int nblk = 256;
  double *part = P.part + 2L * P.nblocks * 4;
  if (ke) {
    hipLaunchKernelGGL(k_ke_part, dim3(nblk), dim3(256), 0, c.stream, m.nx, m.ny, dvol, C, u, v, rho1, rho2, part);
    hipLaunchKernelGGL(k_ke_total, dim3(1), dim3(64), 0, c.stream, nblk, (const double *)part, P.scal + 56);
    COVOF_CHECK_LAUNCH();
  }
  if (dC || du || dv) {
    const dim3 blk(64, 4), grd((m.nx + 2 + 63) / 64, (m.ny + 2 + 3) / 4);
    hipLaunchKernelGGL(k_ke_grad, grd, blk, 0, c.stream, m.nx, m.ny, dvol, C, u, v, rho1, rho2, dC, du, dv);
    COVOF_CHECK_LAUNCH();
  }
  if (ke) {
    COVOF_CHECK(hipMemcpyAsync(P.hscal + 56, P.scal + 56, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    COVOF_CHECK(hipStreamSynchronize(c.stream));
    *ke = P.hscal[56];
  }
  return COVOF_OK;
}

int properties_adjoint(Ctx &c, const double *C, double rho1, double rho2, const double *g, double *dC) {
  const Mesh &m = c.m;
  const dim3 blk(64, 4), grd((m.nx + 63) / 64, (m.ny + 3) / 4);
  hipLaunchKernelGGL(k_props_adj, grd, blk, 0, c.stream, m.nx, m.ny, C, rho1, rho2, g, dC);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int inertial_predictor(Ctx &c, int first, double dt, double gx, double gy, const double *u, const double *v,
                       const double *un1, const double *vn1, double *us, double *vs) {
  if (!closed_or_periodic(c.m)) return COVOF_BAD_ARGUMENT;
  const Mesh &m = c.m;
  const dim3 blk(64, 4), grd((m.nx + 2 + 63) / 64, (m.ny + 2 + 3) / 4);
  hipLaunchKernelGGL(k_pred, grd, blk, 0, c.stream, pred_faces(m), first, dt, gx, gy, u, v, un1, vn1, us, vs);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int inertial_predictor_adjoint(Ctx &c, int first, const double *wu, const double *wv, double *du, double *dv,
                               double *dun1, double *dvn1) {
  if (!closed_or_periodic(c.m)) return COVOF_BAD_ARGUMENT;
  const Mesh &m = c.m;
  const dim3 blk(64, 4), grd((m.nx + 2 + 63) / 64, (m.ny + 2 + 3) / 4);
  hipLaunchKernelGGL(k_pred_adj, grd, blk, 0, c.stream, pred_faces(m), first, wu, wv, du, dv, dun1, dvn1);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

}  // namespace covof
