// The momentum block of the standard projection (the SP velocity predictor):
// src/module_simulation.f90:2454-2469 —
//   assemble_preconditioner_blocks (the Jacobi diagonal)
//       src/module_navier_stokes.f90:1268-1595
//   assemble_rhs                        src/module_navier_stokes.f90:785-1101
//   pack_vector / unpack_vector         src/module_navier_stokes.f90:231-375
//   apply_system_matrix                 src/module_navier_stokes.f90:380-780
//   solve_linear_system (BiCGSTAB(2), CG; NONE / JACOBI)
//       src/module_linear_solvers.f90:88-135, :361-579, :735-838, :1167-1208
//   sync_u / sync_v + apply_velocity_bc src/module_mpi_comm.f90:46-56, :290-537
//
// Supported subset (everything else returns COVOF_BAD_ARGUMENT at bind time):
// the standard projection (no augmented-Lagrangian or monolithic terms), no
// obstacles, UNIFORM boundary profiles, no energy equation, no SSP-RK3
// momentum, and no NEUMANN top side (the reference's v-row at j = ny+1 then
// reads past its array).
//
// Layout: the unknown vector is u(1:u_nx, 1:ny) then v(1:nx, 1:v_ny), i
// fastest (u_nx = nx+1 and v_ny = ny+1 on physical sides), device FP64.
// Every kernel evaluates the reference's expressions in its operand order
// (-ffp-contract=off), so diagonal, right-hand side and products are bitwise
// the reference's; the Krylov loops differ from it only in the summation
// order of their inner products.
//
// The operator keeps the reference's persistent work fields (its SAVEd `tf`):
// each product unpacks x into them, exchanges the periodic halos and applies
// the homogeneous velocity BC in the reference's order, and ghost cells that
// sequence does not write keep what the previous product left (zero at first
// use), exactly as the reference's do.
#include "covof_internal.h"
#include "seq_sweep.h"

#include <cmath>

namespace covof {

namespace {

constexpr int BLK = 256;

// dof -> (i, j, is_v)
struct Dof {
  int i, j, isv;
};
__device__ __forceinline__ Dof dof_ij(long k, long nu, int u_nx, int nx) {
  Dof d;
  if (k < nu) {
    d.isv = 0;
    d.j = (int)(k / u_nx) + 1;
    d.i = (int)(k - (long)(d.j - 1) * u_nx) + 1;
  } else {
    const long kv = k - nu;
    d.isv = 1;
    d.j = (int)(kv / nx) + 1;
    d.i = (int)(kv - (long)(d.j - 1) * nx) + 1;
  }
  return d;
}

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ bool wallbc(int t) { return t == BC_DIRICHLET || t == BC_FREESLIP; }

#define HX(i, j) ((long)(j) * (m.nx + 2) + (i))

// conv_face_flux (src/module_navier_stokes.f90:2388-2416): centered, upwind,
// or the superbee-limited TVD flux.
__device__ double conv_face_flux(int scheme, double flux, double phi_L, double phi_R, double q_LL, double q_L,
                                 double q_R, double q_RR, bool has_LL, bool has_RR) {
  if (scheme == 2) return 0.5 * flux * (phi_L + phi_R);
  if (scheme == 1) return flux >= 0.0 ? flux * phi_L : flux * phi_R;
  const double dq = q_R - q_L;
  double psi = 0.0;
  if (flux >= 0.0) {
    if (has_LL && fabs(dq) > 1.0e-14) {
      const double r = (q_L - q_LL) / dq;
      psi = dmax(dmax(0.0, dmin(2.0 * r, 1.0)), dmin(r, 2.0));
    }
    return flux * (phi_L + 0.5 * psi * (phi_R - phi_L));
  }
  if (has_RR && fabs(dq) > 1.0e-14) {
    const double r = (q_RR - q_R) / dq;
    psi = dmax(dmax(0.0, dmin(2.0 * r, 1.0)), dmin(r, 2.0));
  }
  return flux * (phi_R + 0.5 * psi * (phi_L - phi_R));
}

// compute_u_conv_fluxes (src/module_navier_stokes.f90:2418-2457)
__device__ void u_fluxes(const Mesh &m, const MomCfg &c, const MomFields &F, int i, int j, double &FE,
                         double &FW, double &FN, double &FS) {
  const double *uu = c.lin == 2 ? F.uo : F.u, *vv = c.lin == 2 ? F.vo : F.v;
  double u_n = uu[HX(i, j)], u_n_E = uu[HX(i + 1, j)], u_n_W = uu[HX(i - 1, j)];
  double v_n_N = 0.5 * (vv[HX(i - 1, j + 1)] + vv[HX(i, j + 1)]);
  double v_n_S = 0.5 * (vv[HX(i - 1, j)] + vv[HX(i, j)]);
  if (c.ab_nm1) {
    u_n = 2.0 * u_n - F.un[HX(i, j)];
    u_n_E = 2.0 * u_n_E - F.un[HX(i + 1, j)];
    u_n_W = 2.0 * u_n_W - F.un[HX(i - 1, j)];
    const double v_nm1_N = 0.5 * (F.vn[HX(i - 1, j + 1)] + F.vn[HX(i, j + 1)]);
    const double v_nm1_S = 0.5 * (F.vn[HX(i - 1, j)] + F.vn[HX(i, j)]);
    v_n_N = 2.0 * v_n_N - v_nm1_N;
    v_n_S = 2.0 * v_n_S - v_nm1_S;
  }
  FE = 0.5 * (u_n + u_n_E) * m.dy;
  FW = 0.5 * (u_n_W + u_n) * m.dy;
  FN = v_n_N * m.dx;
  FS = v_n_S * m.dx;
}

// compute_v_conv_fluxes (src/module_navier_stokes.f90:2459-2498)
__device__ void v_fluxes(const Mesh &m, const MomCfg &c, const MomFields &F, int i, int j, double &FE,
                         double &FW, double &FN, double &FS) {
  const double *uu = c.lin == 2 ? F.uo : F.u, *vv = c.lin == 2 ? F.vo : F.v;
  double v_n = vv[HX(i, j)], v_n_N = vv[HX(i, j + 1)], v_n_S = vv[HX(i, j - 1)];
  double u_n_E = 0.5 * (uu[HX(i + 1, j - 1)] + uu[HX(i + 1, j)]);
  double u_n_W = 0.5 * (uu[HX(i, j - 1)] + uu[HX(i, j)]);
  if (c.ab_nm1) {
    v_n = 2.0 * v_n - F.vn[HX(i, j)];
    v_n_N = 2.0 * v_n_N - F.vn[HX(i, j + 1)];
    v_n_S = 2.0 * v_n_S - F.vn[HX(i, j - 1)];
    const double u_nm1_E = 0.5 * (F.un[HX(i + 1, j - 1)] + F.un[HX(i + 1, j)]);
    const double u_nm1_W = 0.5 * (F.un[HX(i, j - 1)] + F.un[HX(i, j)]);
    u_n_E = 2.0 * u_n_E - u_nm1_E;
    u_n_W = 2.0 * u_n_W - u_nm1_W;
  }
  FE = u_n_E * m.dy;
  FW = u_n_W * m.dy;
  FN = 0.5 * (v_n + v_n_N) * m.dx;
  FS = 0.5 * (v_n_S + v_n) * m.dx;
}

// x -> interior unknowns of the halo fields u, v (and p beyond nuv: the
// monolithic system) -- ns_unpack_vector's copy loops
__global__ void __launch_bounds__(BLK) k_unpack(Mesh m, long nu, long nuv, long n, int u_nx,
                                                const double *__restrict__ x, double *__restrict__ u,
                                                double *__restrict__ v, double *__restrict__ p) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  if (k >= nuv) {
    const long kp = k - nuv;
    const int j = (int)(kp / m.nx) + 1;
    p[HX((int)(kp - (long)(j - 1) * m.nx) + 1, j)] = x[k];
    return;
  }
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  (d.isv ? v : u)[HX(d.i, d.j)] = x[k];
}

// ns_pack_vector
__global__ void __launch_bounds__(BLK) k_pack(Mesh m, long nu, long nuv, long n, int u_nx,
                                              const double *__restrict__ u, const double *__restrict__ v,
                                              const double *__restrict__ p, double *__restrict__ x) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  if (k >= nuv) {
    const long kp = k - nuv;
    const int j = (int)(kp / m.nx) + 1;
    x[k] = p[HX((int)(kp - (long)(j - 1) * m.nx) + 1, j)];
    return;
  }
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  x[k] = (d.isv ? v : u)[HX(d.i, d.j)];
}

// sync_u, sync_v (exchange_halos_scalar on one rank: periodic columns over
// all rows, then periodic rows over all columns) and apply_velocity_bc, in
// the reference's statement order: one workgroup, a barrier between every
// step whose reads depend on an earlier step's writes.
// The single-rank exchange_halos_scalar of one halo field inside a
// workgroup (all threads of the block call it).
__device__ void wg_exchange(const Mesh &m, double *f) {
  const int nx = m.nx, ny = m.ny, t = threadIdx.x, T = blockDim.x;
  if (m.perx) {
    for (int j = t; j <= ny + 1; j += T) f[HX(0, j)] = f[HX(nx, j)];
    __syncthreads();
    for (int j = t; j <= ny + 1; j += T) f[HX(nx + 1, j)] = f[HX(1, j)];
    __syncthreads();
  }
  if (m.pery) {
    for (int i = t; i <= nx + 1; i += T) f[HX(i, 0)] = f[HX(i, ny)];
    __syncthreads();
    for (int i = t; i <= nx + 1; i += T) f[HX(i, ny + 1)] = f[HX(i, 1)];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(1024) k_vel_halo(Mesh m, MomCfg c, double *__restrict__ u,
                                                   double *__restrict__ v, int homogeneous, int exchange,
                                                   double *__restrict__ p) {
  const int nx = m.nx, ny = m.ny, t = threadIdx.x, T = blockDim.x;
  const int *bc = m.bc;
  const bool bl = !m.perx, bb = !m.pery;  // physical sides (single rank)
  if (exchange) {
    if (p) wg_exchange(m, p);  // sync_cell_scalar of the monolithic pressure
    wg_exchange(m, u);
    wg_exchange(m, v);
  }
  // apply_velocity_bc (src/module_mpi_comm.f90:290-537), UNIFORM profiles
  const double u_l = homogeneous ? 0.0 : c.bcu[0], v_l = homogeneous ? 0.0 : c.bcv[0];
  const double u_r = homogeneous ? 0.0 : c.bcu[1], v_r = homogeneous ? 0.0 : c.bcv[1];
  const double u_b = homogeneous ? 0.0 : c.bcu[2], v_b = homogeneous ? 0.0 : c.bcv[2];
  const double u_t = homogeneous ? 0.0 : c.bcu[3], v_t = homogeneous ? 0.0 : c.bcv[3];
  if (bc[0] == BC_PERIODIC) {
    for (int j = 1 + t; j <= ny; j += T) {
      u[HX(0, j)] = u[HX(nx, j)];
      u[HX(nx + 1, j)] = u[HX(1, j)];
      v[HX(0, j)] = v[HX(nx, j)];
      v[HX(nx + 1, j)] = v[HX(1, j)];
    }
    __syncthreads();
  }
  if (bc[2] == BC_PERIODIC) {
    for (int i = 1 + t; i <= nx; i += T) {
      v[HX(i, 0)] = v[HX(i, ny)];
      v[HX(i, ny + 1)] = v[HX(i, 1)];
    }
    __syncthreads();
    for (int i = t; i <= nx + 1; i += T) {
      u[HX(i, 0)] = u[HX(i, ny)];
      u[HX(i, ny + 1)] = u[HX(i, 1)];
    }
    __syncthreads();
  }
  if (bl && bc[0] != BC_PERIODIC) {
    for (int j = 1 + t; j <= ny; j += T) {
      if (bc[0] == BC_DIRICHLET) {
        if (!homogeneous) u[HX(1, j)] = u_l;
        u[HX(0, j)] = u_l;
        v[HX(0, j)] = 2.0 * v_l - v[HX(1, j)];
      } else if (bc[0] == BC_FREESLIP) {
        if (!homogeneous) u[HX(1, j)] = 0.0;
        u[HX(0, j)] = 0.0;
        v[HX(0, j)] = v[HX(1, j)];
      } else if (bc[0] == BC_NEUMANN) {
        u[HX(0, j)] = u[HX(1, j)];
        v[HX(0, j)] = v[HX(1, j)];
      }
    }
    __syncthreads();
  }
  if (bl && bc[1] != BC_PERIODIC) {
    for (int j = 1 + t; j <= ny; j += T) {
      if (bc[1] == BC_DIRICHLET) {
        if (!homogeneous) u[HX(nx + 1, j)] = u_r;
        v[HX(nx + 1, j)] = 2.0 * v_r - v[HX(nx, j)];
      } else if (bc[1] == BC_FREESLIP) {
        if (!homogeneous) u[HX(nx + 1, j)] = 0.0;
        v[HX(nx + 1, j)] = v[HX(nx, j)];
      } else if (bc[1] == BC_NEUMANN) {
        if (!homogeneous) u[HX(nx + 1, j)] = u[HX(nx, j)];
        v[HX(nx + 1, j)] = v[HX(nx, j)];
      }
    }
    __syncthreads();
  }
  if (bb && bc[2] != BC_PERIODIC) {
    for (int i = 1 + t; i <= nx; i += T) {
      if (bc[2] == BC_DIRICHLET) {
        if (!homogeneous) v[HX(i, 1)] = v_b;
        v[HX(i, 0)] = v_b;
        if (ny >= 2) u[HX(i, 0)] = (8.0 * u_b - 6.0 * u[HX(i, 1)] + u[HX(i, 2)]) / 3.0;
        else u[HX(i, 0)] = 2.0 * u_b - u[HX(i, 1)];
      } else if (bc[2] == BC_FREESLIP) {
        if (!homogeneous) v[HX(i, 1)] = 0.0;
        v[HX(i, 0)] = 0.0;
        u[HX(i, 0)] = u[HX(i, 1)];
      } else if (bc[2] == BC_NEUMANN) {
        u[HX(i, 0)] = u[HX(i, 1)];
        v[HX(i, 0)] = v[HX(i, 1)];
      }
    }
    __syncthreads();
  }
  if (bb && bc[3] != BC_PERIODIC) {
    for (int i = 1 + t; i <= nx; i += T) {
      if (bc[3] == BC_DIRICHLET) {
        if (!homogeneous) v[HX(i, ny + 1)] = v_t;
        if (ny >= 2) u[HX(i, ny + 1)] = (8.0 * u_t - 6.0 * u[HX(i, ny)] + u[HX(i, ny - 1)]) / 3.0;
        else u[HX(i, ny + 1)] = 2.0 * u_t - u[HX(i, ny)];
      } else if (bc[3] == BC_FREESLIP) {
        if (!homogeneous) v[HX(i, ny + 1)] = 0.0;
        u[HX(i, ny + 1)] = u[HX(i, ny)];
      } else if (bc[3] == BC_NEUMANN) {
        u[HX(i, ny + 1)] = u[HX(i, ny)];
        if (!homogeneous) v[HX(i, ny + 1)] = v[HX(i, ny)];
      }
    }
    __syncthreads();
  }
  if (!homogeneous && t == 0) {
    if (bl && bc[1] != BC_PERIODIC) {
      if (bb && bc[2] != BC_PERIODIC) {
        u[HX(nx + 1, 0)] = u[HX(nx + 1, 1)];
        v[HX(nx + 1, 0)] = v[HX(nx, 0)];
      }
      if (bb && bc[3] != BC_PERIODIC) {
        u[HX(nx + 1, ny + 1)] = u[HX(nx + 1, ny)];
        v[HX(nx + 1, ny + 1)] = v[HX(nx, ny + 1)];
      }
    }
    if (bl && bc[0] != BC_PERIODIC && bb && bc[3] != BC_PERIODIC) v[HX(0, ny + 1)] = v[HX(1, ny + 1)];
  }
}

// ns_apply_system_matrix, SP branch (src/module_navier_stokes.f90:380-780):
// row k of A x from the unpacked work fields tu, tv.
__device__ __forceinline__ double mom_row(const Mesh &m, const MomCfg &c, const MomFields &F, double dt, long k,
                                          long nu, int u_nx, int v_ny, const double *__restrict__ tu,
                                          const double *__restrict__ tv, const double *__restrict__ tp) {
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  const int i = d.i, j = d.j, nx = m.nx, ny = m.ny;
  const double dx = m.dx, dy = m.dy, dvol = dx * dy;
  const double at = c.gear ? 1.5 : 1.0;
  const bool bl = !m.perx, bb = !m.pery;
  const int sch = c.scheme;
  const bool has_ee = i + 2 <= nx + 1, has_ww = i - 2 >= 0, has_nn = j + 2 <= ny + 1, has_ss = j - 2 >= 0;
#define TU(a, b) tu[HX(a, b)]
#define TV(a, b) tv[HX(a, b)]
  if (!d.isv) {
    const double *u = F.u;
    const double rho_u = F.ru[HX(i, j)];
    const double diag_bc = (at * rho_u / dt) * dvol;
    if (i == 1 && bl) {
      return m.bc[0] == BC_NEUMANN ? diag_bc * (TU(1, j) - TU(2, j)) : diag_bc * TU(1, j);
    }
    if (i == u_nx && bl) {
      return m.bc[1] == BC_NEUMANN ? diag_bc * (TU(u_nx, j) - TU(u_nx - 1, j)) : diag_bc * TU(u_nx, j);
    }
    // every operand is loaded before the row is formed, without a condition
    // (a stencil cell the row does not have reads the row's own cell and is
    // not used), so the loads are in flight together
    const double tuP = TU(i, j), tuE = TU(i + 1, j), tuW = TU(i - 1, j), tuN = TU(i, j + 1), tuS = TU(i, j - 1);
    const double tvSW = TV(i - 1, j), tvSE = TV(i, j), tvNW = TV(i - 1, j + 1), tvNE = TV(i, j + 1);
    const double mu_E = F.mu[HX(i, j)], mu_W = F.mu[HX(i - 1, j)];
    const double mnN = F.mn[HX(i, j + 1)], mnS = F.mn[HX(i, j)];
    const double uW = u[HX(i - 1, j)], uP = u[HX(i, j)], uE = u[HX(i + 1, j)];
    const double uS = u[HX(i, j - 1)], uN = u[HX(i, j + 1)];
    const double r_ee = u[HX(has_ee ? i + 2 : i, j)], r_ww = u[HX(has_ww ? i - 2 : i, j)];
    const double r_nn = u[HX(i, has_nn ? j + 2 : j)], r_ss = u[HX(i, has_ss ? j - 2 : j)];
    double val = (at * rho_u / dt) * tuP * dvol;
    if (c.conv) {
      double cE, cW, cN, cS;
      u_fluxes(m, c, F, i, j, cE, cW, cN, cS);
      const double q_ee = has_ee ? r_ee : 0.0, q_ww = has_ww ? r_ww : 0.0;
      const double q_nn = has_nn ? r_nn : 0.0, q_ss = has_ss ? r_ss : 0.0;
      const double pE = conv_face_flux(sch, cE, tuP, tuE, uW, uP, uE, q_ee, true, has_ee);
      const double pW = conv_face_flux(sch, cW, tuW, tuP, q_ww, uW, uP, uE, has_ww, true);
      const double pN = conv_face_flux(sch, cN, tuP, tuN, uS, uP, uN, q_nn, true, has_nn);
      const double pS = conv_face_flux(sch, cS, tuS, tuP, q_ss, uS, uP, uN, has_ss, true);
      val = val + rho_u * (pE - pW + pN - pS);
      if (c.lin == 3) {
        const double du_dx = (uE - uW) / (2.0 * dx);
        const double du_dy = (uN - uS) / (2.0 * dy);
        const double tvf = 0.25 * (tvSW + tvSE + tvNW + tvNE);
        val = val + rho_u * (tuP * du_dx + tvf * du_dy) * dvol;
      }
    }
    const double deE = 2.0 * mu_E * (tuE - tuP) / dx * dy;
    const double deW = 2.0 * mu_W * (tuP - tuW) / dx * dy;
    val = val - (deE - deW);
    const double shN = mnN * ((tuN - tuP) / dy + (tvNE - tvNW) / dx) * dx;
    const double shS = mnS * ((tuP - tuS) / dy + (tvSE - tvSW) / dx) * dx;
    val = val - (shN - shS);
    if (c.mono) val = val + (tp[HX(i, j)] - tp[HX(i - 1, j)]) * dy;  // B_u^T p
    return val;
  } else {
    const double *v = F.v;
    const double rho_v = F.rv[HX(i, j)];
    const double diag_bc = (at * rho_v / dt) * dvol;
    if ((j == 1 && bb && wallbc(m.bc[2])) || (j == v_ny && bb && wallbc(m.bc[3]))) {
      return diag_bc * TV(i, j);
    }
    const double tvP = TV(i, j), tvE = TV(i + 1, j), tvW = TV(i - 1, j), tvN = TV(i, j + 1), tvS = TV(i, j - 1);
    const double tuSW = TU(i, j - 1), tuSE = TU(i + 1, j - 1), tuNW = TU(i, j), tuNE = TU(i + 1, j);
    const double mu_N = F.mu[HX(i, j)], mu_S = F.mu[HX(i, j - 1)];
    const double mnE = F.mn[HX(i + 1, j)], mnW = F.mn[HX(i, j)];
    const double vW = v[HX(i - 1, j)], vP = v[HX(i, j)], vE = v[HX(i + 1, j)];
    const double vS = v[HX(i, j - 1)], vN = v[HX(i, j + 1)];
    const double r_ee = v[HX(has_ee ? i + 2 : i, j)], r_ww = v[HX(has_ww ? i - 2 : i, j)];
    const double r_nn = v[HX(i, has_nn ? j + 2 : j)], r_ss = v[HX(i, has_ss ? j - 2 : j)];
    double val = (at * rho_v / dt) * tvP * dvol;
    if (c.conv) {
      double cE, cW, cN, cS;
      v_fluxes(m, c, F, i, j, cE, cW, cN, cS);
      const double q_ee = has_ee ? r_ee : 0.0, q_ww = has_ww ? r_ww : 0.0;
      const double q_nn = has_nn ? r_nn : 0.0, q_ss = has_ss ? r_ss : 0.0;
      const double pE = conv_face_flux(sch, cE, tvP, tvE, vW, vP, vE, q_ee, true, has_ee);
      const double pW = conv_face_flux(sch, cW, tvW, tvP, q_ww, vW, vP, vE, has_ww, true);
      const double pN = conv_face_flux(sch, cN, tvP, tvN, vS, vP, vN, q_nn, true, has_nn);
      const double pS = conv_face_flux(sch, cS, tvS, tvP, q_ss, vS, vP, vN, has_ss, true);
      val = val + rho_v * (pE - pW + pN - pS);
      if (c.lin == 3) {
        const double dv_dx = (vE - vW) / (2.0 * dx);
        const double dv_dy = (vN - vS) / (2.0 * dy);
        const double tuf = 0.25 * (tuSW + tuSE + tuNW + tuNE);
        val = val + rho_v * (tuf * dv_dx + tvP * dv_dy) * dvol;
      }
    }
    const double deN = 2.0 * mu_N * (tvN - tvP) / dy * dx;
    const double deS = 2.0 * mu_S * (tvP - tvS) / dy * dx;
    val = val - (deN - deS);
    const double shE = mnE * ((tvE - tvP) / dx + (tuNE - tuSE) / dy) * dy;
    const double shW = mnW * ((tvP - tvW) / dx + (tuNW - tuSW) / dy) * dy;
    val = val - (shE - shW);
    if (c.mono) val = val + (tp[HX(i, j)] - tp[HX(i, j - 1)]) * dx;  // B_v^T p
    return val;
  }
#undef TU
#undef TV
}

// y = A x, one unknown per thread; with JAC, y = M^-1 A x for the Jacobi
// preconditioner (apply_PA, src/module_linear_solvers.f90:323-356): the same
// row value divided as apply_preconditioner divides it.
template <bool JAC>
__global__ void __launch_bounds__(BLK) k_mom_matvec(Mesh m, MomCfg c, MomFields F, double dt, long nu, long nuv,
                                                    long n, int u_nx, int v_ny, const double *__restrict__ tu,
                                                    const double *__restrict__ tv, const double *__restrict__ tp,
                                                    const double *__restrict__ diag, double *__restrict__ y) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  double val;
  if (k >= nuv) {
    // continuity row of the monolithic system: div(u) dvol, and the
    // regularisation 1e-10 p(1,1) dvol at cell (1,1)
    const long kp = k - nuv;
    const int j = (int)(kp / m.nx) + 1, i = (int)(kp - (long)(j - 1) * m.nx) + 1;
    const double dvol = m.dx * m.dy;
    const double div = (tu[HX(i + 1, j)] - tu[HX(i, j)]) / m.dx + (tv[HX(i, j + 1)] - tv[HX(i, j)]) / m.dy;
    val = (i == 1 && j == 1) ? div * dvol + 1.0e-10 * tp[HX(1, 1)] * dvol : div * dvol;
  } else {
    val = mom_row(m, c, F, dt, k, nu, u_nx, v_ny, tu, tv, tp);
  }
  if (JAC) {
    const double dk = diag[k];
    y[k] = fabs(dk) > 1.0e-15 ? val / dk : val;
  } else {
    y[k] = val;
  }
}

// Outputs of the blocks kernel; any pointer may be null.  u*: (u_nx, ny),
// v*: (nx, v_ny) block coefficients; uaP2, vaP2: u_aP, v_aP on the halo
// grid, the Schur coefficients' source.
struct BlkOut {
  double *diag;
  double *uW, *uE, *uS, *uN, *uP;
  double *vW, *vE, *vS, *vN, *vP;
  double *uaP2, *vaP2;
};
#define PUT(arr, val)                    \
  do {                                   \
    if (o.arr) o.arr[dof] = (val);       \
  } while (0)

// ns_assemble_preconditioner_blocks (src/module_navier_stokes.f90:1268-1595;
// no AL, Brinkman or obstacle terms) over the u and v unknowns: diag_u,
// diag_v, the SSOR coefficients u_a*, v_a*, and u_aP / v_aP on the halo grid.
__global__ void __launch_bounds__(BLK) k_mom_blocks(Mesh m, MomCfg c, MomFields F, double dt, long nu, long nuv,
                                                    int u_nx, int v_ny, BlkOut o) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= nuv) return;
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  const int i = d.i, j = d.j, nx = m.nx, ny = m.ny;
  const double dx = m.dx, dy = m.dy, dvol = dx * dy;
  const double dt_safe = dt > 1.0e-30 ? dt : 1.0;
  const double at = c.gear ? 1.5 : 1.0;
  const bool bl = !m.perx, bb = !m.pery;
  if (!d.isv) {
    const long dof = k;
    const double rho_u = F.ru[HX(i, j)];
    double Du = (at * rho_u / dt_safe) * dvol;
    if (Du <= 1.0e-30) Du = 1.0;
    if ((i == 1 && bl) || (i == u_nx && bl)) {
      const bool west = i == 1;
      if (o.diag) o.diag[k] = Du;
      if (o.uaP2) o.uaP2[HX(i, j)] = Du;
      const double nb = (m.bc[west ? 0 : 1] == BC_NEUMANN && u_nx >= 2) ? Du : 0.0;
      PUT(uW, west ? 0.0 : nb);
      PUT(uE, west ? nb : 0.0);
      PUT(uS, 0.0);
      PUT(uN, 0.0);
      PUT(uP, Du);
      return;
    }
    const double mu_e = F.mu[HX(i, j)], mu_w = F.mu[HX(i - 1, j)];
    const double cE = 2.0 * mu_e * (dy / dx), cW = 2.0 * mu_w * (dy / dx);
    const double cN_raw = F.mn[HX(i, j + 1)] * (dx / dy), cS_raw = F.mn[HX(i, j)] * (dx / dy);
    double cN_visc = cN_raw, cS_visc = cS_raw, dwy = 0.0, cN_extra = 0.0, cS_extra = 0.0;
    if (j == 1 && bb) {
      if (m.bc[2] == BC_DIRICHLET) {
        if (ny >= 2) {
          dwy = dwy + 3.0 * cS_raw;
          cN_extra = cN_extra + cS_raw / 3.0;
        } else {
          dwy = dwy + 2.0 * cS_raw;
        }
      }
      cS_visc = 0.0;
    }
    if (j == ny && bb) {
      if (m.bc[3] == BC_DIRICHLET) {
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
    if (c.conv) {
      double FE, FW, FN, FS;
      u_fluxes(m, c, F, i, j, FE, FW, FN, FS);
      conv_diag = rho_u * (dmax(FE, 0.0) + dmax(-FW, 0.0) + dmax(FN, 0.0) + dmax(-FS, 0.0));
      conv_E = rho_u * dmax(-FE, 0.0);
      conv_W = rho_u * dmax(FW, 0.0);
      conv_N = (j == ny && bb) ? 0.0 : rho_u * dmax(-FN, 0.0);
      conv_S = (j == 1 && bb) ? 0.0 : rho_u * dmax(FS, 0.0);
      if (c.lin == 3) {
        const double du_dx = (F.u[HX(i + 1, j)] - F.u[HX(i - 1, j)]) / (2.0 * dx);
        newt = rho_u * dmax(0.0, du_dx) * dvol;
      }
    }
    PUT(uW, cW + conv_W + 0.0);
    PUT(uE, cE + conv_E + 0.0);
    PUT(uS, cS_visc + conv_S + cS_extra);
    PUT(uN, cN_visc + conv_N + cN_extra);
    double aP = Du + dwy + conv_diag + 0.0 + newt + cW + cE + cS_visc + cN_visc + 0.0 + 0.0;
    if (aP <= 1.0e-15) aP = 1.0;
    if (o.diag) o.diag[k] = aP;
    PUT(uP, aP);
    if (o.uaP2) o.uaP2[HX(i, j)] = aP;
  } else {
    const long dof = k - nu;
    const double rho_v = F.rv[HX(i, j)];
    double Dv = (at * rho_v / dt_safe) * dvol;
    if (Dv <= 1.0e-30) Dv = 1.0;
    const bool is_wall = (j == 1 && bb && wallbc(m.bc[2])) || (j == v_ny && bb && wallbc(m.bc[3]));
    const bool nm_s = j == 1 && bb && m.bc[2] == BC_NEUMANN && v_ny >= 2;
    const bool nm_n = j == v_ny && bb && m.bc[3] == BC_NEUMANN && v_ny >= 2;
    if (is_wall || nm_s || nm_n) {
      if (o.diag) o.diag[k] = Dv;
      if (o.vaP2) o.vaP2[HX(i, j)] = Dv;
      PUT(vE, 0.0);
      PUT(vW, 0.0);
      PUT(vN, (!is_wall && nm_s) ? Dv : 0.0);
      PUT(vS, (!is_wall && !nm_s && nm_n) ? Dv : 0.0);
      PUT(vP, Dv);
      return;
    }
    const double mu_n = F.mu[HX(i, j)], mu_s = F.mu[HX(i, j - 1)];
    const double cN = 2.0 * mu_n * (dx / dy), cS = 2.0 * mu_s * (dx / dy);
    const double cE_raw = F.mn[HX(i + 1, j)] * (dy / dx), cW_raw = F.mn[HX(i, j)] * (dy / dx);
    double cE_visc = cE_raw, cW_visc = cW_raw, dwx = 0.0;
    if (i == 1 && bl) {
      if (m.bc[0] == BC_DIRICHLET) dwx = dwx + 2.0 * cW_raw;
      cW_visc = 0.0;
    }
    if (i == nx && bl) {
      if (m.bc[1] == BC_DIRICHLET) dwx = dwx + 2.0 * cE_raw;
      cE_visc = 0.0;
    }
    double conv_diag = 0.0, conv_W = 0.0, conv_E = 0.0, conv_S = 0.0, conv_N = 0.0, newt = 0.0;
    if (c.conv) {
      double FE, FW, FN, FS;
      v_fluxes(m, c, F, i, j, FE, FW, FN, FS);
      conv_diag = rho_v * (dmax(FE, 0.0) + dmax(-FW, 0.0) + dmax(FN, 0.0) + dmax(-FS, 0.0));
      conv_E = (i == nx && bl) ? 0.0 : rho_v * dmax(-FE, 0.0);
      conv_W = (i == 1 && bl) ? 0.0 : rho_v * dmax(FW, 0.0);
      conv_N = rho_v * dmax(-FN, 0.0);
      conv_S = rho_v * dmax(FS, 0.0);
      if (c.lin == 3) {
        const double dv_dy = (F.v[HX(i, j + 1)] - F.v[HX(i, j - 1)]) / (2.0 * dy);
        newt = rho_v * dmax(0.0, dv_dy) * dvol;
      }
    }
    PUT(vW, cW_visc + conv_W);
    PUT(vE, cE_visc + conv_E);
    PUT(vS, cS + conv_S + 0.0);
    PUT(vN, cN + conv_N + 0.0);
    double aP = Dv + dwx + conv_diag + 0.0 + newt + cW_visc + cE_visc + cS + cN + 0.0 + 0.0;
    if (aP <= 1.0e-15) aP = 1.0;
    if (o.diag) o.diag[k] = aP;
    PUT(vP, aP);
    if (o.vaP2) o.vaP2[HX(i, j)] = aP;
  }
}
#undef PUT

// The Schur block of the monolithic system (same routine, :1540-1590): the
// 5-point coefficients of B diag(A)^-1 B^T from u_aP, v_aP (halo, exchanged),
// no coupling through wall sides, and the regularisation (1e-10 dvol at cell
// (1,1), 1e-12 dvol elsewhere).  diag: the pressure rows of diag_p.
__global__ void __launch_bounds__(BLK) k_schur(Mesh m, const double *__restrict__ uaP2,
                                               const double *__restrict__ vaP2, double *__restrict__ sW,
                                               double *__restrict__ sE, double *__restrict__ sS,
                                               double *__restrict__ sN, double *__restrict__ sP,
                                               double *__restrict__ diag) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= (long)m.nx * m.ny) return;
  const int j = (int)(k / m.nx) + 1, i = (int)(k - (long)(j - 1) * m.nx) + 1;
  const bool bl = !m.perx, bb = !m.pery;
  const double dx = m.dx, dy = m.dy, dvol = dx * dy;
  double aW = 0.0, aE = 0.0, aS = 0.0, aN = 0.0;
  if (!(i == 1 && bl && wallbc(m.bc[0]))) aW = (dy * dy) / dmax(1.0e-30, uaP2[HX(i, j)]);
  if (!(i == m.nx && bl && wallbc(m.bc[1]))) aE = (dy * dy) / dmax(1.0e-30, uaP2[HX(i + 1, j)]);
  if (!(j == 1 && bb && wallbc(m.bc[2]))) aS = (dx * dx) / dmax(1.0e-30, vaP2[HX(i, j)]);
  if (!(j == m.ny && bb && wallbc(m.bc[3]))) aN = (dx * dx) / dmax(1.0e-30, vaP2[HX(i, j + 1)]);
  const double reg = (i == 1 && j == 1) ? 1.0e-10 * dvol : 1.0e-12 * dvol;
  const double aP = aW + aE + aS + aN + 0.0 + reg;
  if (sW) {
    sW[k] = aW;
    sE[k] = aE;
    sS[k] = aS;
    sN[k] = aN;
    sP[k] = aP;
  }
  if (diag) diag[k] = aP;
}

__global__ void __launch_bounds__(BLK) k_fill(long n, double val, double *__restrict__ a) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k < n) a[k] = val;
}

// exchange_halos_scalar of up to two halo fields, one workgroup
__global__ void __launch_bounds__(1024) k_exchange2(Mesh m, double *a, double *b) {
  if (a) wg_exchange(m, a);
  if (b) wg_exchange(m, b);
}

// natural (dof = (j-1) ni + i-1) -> field f of the skewed records
__global__ void __launch_bounds__(BLK) k_to_skew(int ni, int nj, int S, const double *__restrict__ a,
                                                 double *__restrict__ rec, int f) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= (long)ni * nj) return;
  const int j = (int)(k / ni) + 1, i = (int)(k - (long)(j - 1) * ni) + 1;
  rec[sk_of(S, i, j) * 8 + f] = a[k];
}

// The right-hand side of a block, skewed: the residual, or (kind 1: u, 2: v)
// the residual less the pressure gradient of the block-triangular step (xp: a
// halo field).
__global__ void __launch_bounds__(BLK) k_seq_rhs(Mesh m, SeqBlk b, int kind, const double *__restrict__ src,
                                                 const double *__restrict__ xp) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= (long)b.ni * b.nj) return;
  const int j = (int)(k / b.ni) + 1, i = (int)(k - (long)(j - 1) * b.ni) + 1;
  const bool bl = !m.perx, bb = !m.pery;
  double v = src[k];
  if (kind == 1) {
    const bool is_wall = (i == 1 && bl) || (i == b.ni && bl);
    if (!is_wall) v = src[k] - (xp[HX(i, j)] - xp[HX(i - 1, j)]) * m.dy;
  } else if (kind == 2) {
    const bool is_wall = (j == 1 && bb && wallbc(m.bc[2])) || (j == b.nj && bb && wallbc(m.bc[3]));
    if (!is_wall) v = src[k] - (xp[HX(i, j)] - xp[HX(i, j - 1)]) * m.dx;
  }
  b.rec[sk_of(b.S, i, j) * 8 + SR_R] = v;
}

// skewed state -> the block's part of a vector (dst) or the interior of a
// halo field (halo != 0)
__global__ void __launch_bounds__(BLK) k_from_skew(Mesh m, SeqBlk b, const double *__restrict__ x,
                                                   double *__restrict__ dst, int halo) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= (long)b.ni * b.nj) return;
  const int j = (int)(k / b.ni) + 1, i = (int)(k - (long)(j - 1) * b.ni) + 1;
  dst[halo ? HX(i, j) : k] = x[sk_of(b.S, i, j)];
}

struct SeqSet {
  SeqBlk blk[3];  // u, v, Schur
};

// One block per workgroup: blocks first .. first + gridDim.x - 1 of the set,
// n_sweeps sweeps with relaxation omega.
__global__ void __launch_bounds__(64 * SEQ_MAXW) k_seq_ssor(Mesh m, SeqSet a, int first, int n_sweeps,
                                                           double omega) {
  seq_sweeps<SEQ_SSOR>(m.perx, m.pery, a.blk[first + blockIdx.x], n_sweeps, omega);
}

// The pressure step of the block-triangular preconditioner after its 12
// sweeps: xp holds the solution's interior; the mean is removed, the halo
// exchanged, and the pressure part of the result written.  The mean is the
// reference's serial sum in its order (bitwise needs that order): one thread
// adds, while the workgroup stages the next 1024 values into LDS.
__global__ void __launch_bounds__(1024) k_bt_mean(Mesh m, double *__restrict__ xp, double *__restrict__ dst) {
  __shared__ double buf[2][1024];
  __shared__ double mean;
  const long np = (long)m.nx * m.ny;
  auto cell = [&](long k) {
    const int j = (int)(k / m.nx) + 1, i = (int)(k - (long)(j - 1) * m.nx) + 1;
    return HX(i, j);
  };
  const int t = threadIdx.x;
  if (t < np) buf[0][t] = xp[cell(t)];
  __syncthreads();
  double sum = 0.0;
  for (long c0 = 0; c0 < np; c0 += 1024) {
    const long kn = c0 + 1024 + t;
    if (kn < np) buf[((c0 >> 10) + 1) & 1][t] = xp[cell(kn)];
    if (t == 0) {
      const double *b = buf[(c0 >> 10) & 1];
      const int cnt = np - c0 < 1024 ? (int)(np - c0) : 1024;
      for (int q = 0; q < cnt; ++q) sum = sum + b[q];
    }
    __syncthreads();
  }
  if (t == 0) mean = sum / (double)np;
  __syncthreads();
  for (long k = t; k < np; k += blockDim.x) {
    const long h = cell(k);
    const double v = xp[h] - mean;
    xp[h] = v;
    dst[k] = v;
  }
  __syncthreads();
  if (m.perx || m.pery) wg_exchange(m, xp);
}

// ns_assemble_rhs, SP branch (src/module_navier_stokes.f90:785-1101)
// (the continuity rows of the monolithic system, k >= nuv, are zero)
__global__ void __launch_bounds__(BLK) k_mom_rhs(Mesh m, MomCfg c, MomFields F, double dt, long nu, long nuv,
                                                 long n, int u_nx, int v_ny, double *__restrict__ rhs) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  if (k >= nuv) {
    rhs[k] = 0.0;
    return;
  }
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  const int i = d.i, j = d.j, nx = m.nx, ny = m.ny;
  const double dx = m.dx, dy = m.dy, dvol = dx * dy;
  const double at = c.gear ? 1.5 : 1.0;
  const bool bl = !m.perx, bb = !m.pery;
  const double *u = F.u, *v = F.v;
  if (!d.isv) {
    const double rho_u = F.ru[HX(i, j)];
    const double diag_bc = (at * rho_u / dt) * dvol;
    if (i == 1 && bl) {
      rhs[k] = diag_bc * (m.bc[0] == BC_DIRICHLET ? c.bcu[0] : 0.0);
      return;
    }
    if (i == u_nx && bl) {
      rhs[k] = diag_bc * (m.bc[1] == BC_DIRICHLET ? c.bcu[1] : 0.0);
      return;
    }
    double val;
    if (c.gear) {
      const double ui = rho_u * (2.0 * F.uo[HX(i, j)] - 0.5 * F.un[HX(i, j)]);
      val = (ui / dt) * dvol;
    } else {
      val = (rho_u / dt) * F.uo[HX(i, j)] * dvol;
    }
    val = val + rho_u * c.gx * dvol;
    val = val + F.fx[HX(i, j)] * dvol;
    if (c.lin == 3) {
      const double du_dx = (u[HX(i + 1, j)] - u[HX(i - 1, j)]) / (2.0 * dx);
      const double du_dy = (u[HX(i, j + 1)] - u[HX(i, j - 1)]) / (2.0 * dy);
      const double vf = 0.25 * (v[HX(i - 1, j)] + v[HX(i, j)] + v[HX(i - 1, j + 1)] + v[HX(i, j + 1)]);
      val = val + rho_u * (u[HX(i, j)] * du_dx + vf * du_dy) * dvol;
    }
    if (j == ny && bb && m.bc[3] == BC_DIRICHLET) {
      const double mu_N = F.mn[HX(i, j + 1)];
      val = ny >= 2 ? val + mu_N * ((8.0 / 3.0) * c.bcu[3] / dy) * dx : val + mu_N * (2.0 * c.bcu[3] / dy) * dx;
    }
    if (j == 1 && bb && m.bc[2] == BC_DIRICHLET) {
      const double mu_S = F.mn[HX(i, j)];
      val = ny >= 2 ? val + mu_S * ((8.0 / 3.0) * c.bcu[2] / dy) * dx : val + mu_S * (2.0 * c.bcu[2] / dy) * dx;
    }
    rhs[k] = val;
  } else {
    const double rho_v = F.rv[HX(i, j)];
    const double diag_bc = (at * rho_v / dt) * dvol;
    if (j == 1 && bb && wallbc(m.bc[2])) {
      rhs[k] = diag_bc * (m.bc[2] == BC_DIRICHLET ? c.bcv[2] : 0.0);
      return;
    }
    if (j == v_ny && bb && wallbc(m.bc[3])) {
      rhs[k] = diag_bc * (m.bc[3] == BC_DIRICHLET ? c.bcv[3] : 0.0);
      return;
    }
    double val;
    if (c.gear) {
      const double vi = rho_v * (2.0 * F.vo[HX(i, j)] - 0.5 * F.vn[HX(i, j)]);
      val = (vi / dt) * dvol;
    } else {
      val = (rho_v / dt) * F.vo[HX(i, j)] * dvol;
    }
    val = val + rho_v * c.gy * dvol;
    val = val + F.fy[HX(i, j)] * dvol;
    if (c.lin == 3) {
      const double dv_dx = (v[HX(i + 1, j)] - v[HX(i - 1, j)]) / (2.0 * dx);
      const double dv_dy = (v[HX(i, j + 1)] - v[HX(i, j - 1)]) / (2.0 * dy);
      const double uf = 0.25 * (u[HX(i, j - 1)] + u[HX(i + 1, j - 1)] + u[HX(i, j)] + u[HX(i + 1, j)]);
      val = val + rho_v * (uf * dv_dx + v[HX(i, j)] * dv_dy) * dvol;
    }
    if (i == 1 && bl && m.bc[0] == BC_DIRICHLET) {
      const double mu_W = F.mn[HX(i, j)];
      val = val + mu_W * (2.0 * c.bcv[0] / dx) * dy;
    }
    if (i == nx && bl && m.bc[1] == BC_DIRICHLET) {
      const double mu_E = F.mn[HX(i + 1, j)];
      val = val + mu_E * (2.0 * c.bcv[1] / dx) * dy;
    }
    rhs[k] = val;
  }
}

// apply_preconditioner JACOBI (src/module_linear_solvers.f90:88-135)
__global__ void __launch_bounds__(BLK) k_mom_jacobi(long n, const double *__restrict__ d,
                                                    const double *__restrict__ r, double *__restrict__ z) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  const double dk = d[k];
  z[k] = fabs(dk) > 1.0e-15 ? r[k] / dk : r[k];
}

#undef HX

inline dim3 grid_of(long n) { return dim3((unsigned)((n + BLK - 1) / BLK)); }

int mom_alloc(Ctx &c) {
  Mom &M = c.mom;
  if (M.tu) return COVOF_OK;
  const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
  // at most: u, v and p unknowns
  const long n = (long)(c.m.nx + 1) * c.m.ny + (long)c.m.nx * (c.m.ny + 1) + (long)c.m.nx * c.m.ny;
  for (double **f : {&M.tu, &M.tv, &M.tp}) {
    COVOF_CHECK(hipMalloc(f, nh * sizeof(double)));
    hipLaunchKernelGGL(k_fill, grid_of(nh), dim3(BLK), 0, c.stream, nh, 0.0, *f);
  }
  COVOF_CHECK_LAUNCH();
  COVOF_CHECK(hipMalloc(&M.diag, n * sizeof(double)));
  COVOF_CHECK(hipMalloc(&M.rhs, n * sizeof(double)));
  COVOF_CHECK(hipMalloc(&M.x, n * sizeof(double)));
  for (double *&w : M.w) COVOF_CHECK(hipMalloc(&w, n * sizeof(double)));
  return COVOF_OK;
}

// the SSOR / Schur coefficients and their work fields, made on first use
int mom_alloc_blocks(Ctx &c) {
  Mom &M = c.mom;
  if (M.bk[0]) return COVOF_OK;
  const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
  const long nb = (long)(c.m.nx + 1) * (c.m.ny + 1);  // covers every block
  for (double *&b : M.bk) COVOF_CHECK(hipMalloc(&b, nb * sizeof(double)));
  for (double *&b : M.aP2) COVOF_CHECK(hipMalloc(&b, nh * sizeof(double)));
  for (double *&b : M.sx) COVOF_CHECK(hipMalloc(&b, nh * sizeof(double)));
  // skewed arrays of the sequential sweeps: u (u_nx, ny), v (nx, v_ny), p (nx, ny)
  const int u_nx = c.m.nx + (c.m.perx ? 0 : 1), v_ny = c.m.ny + (c.m.pery ? 0 : 1);
  const int ni[3] = {u_nx, c.m.nx, c.m.nx}, nj[3] = {c.m.ny, v_ny, c.m.ny};
  for (int b = 0; b < 3; ++b) {
    const long nsk = seq_len(ni[b], nj[b]);
    for (int q = 0; q < 3; ++q) {
      const long len = q == 0 ? 8 * nsk : nsk;  // the records, then the two state buffers
      COVOF_CHECK(hipMalloc(&M.sk[b][q], len * sizeof(double)));
      hipLaunchKernelGGL(k_fill, grid_of(len), dim3(BLK), 0, c.stream, len, 0.0, M.sk[b][q]);
    }
    const long nhal = ni[b] > nj[b] ? ni[b] : nj[b];
    for (double *&p : M.hal[b]) COVOF_CHECK(hipMalloc(&p, nhal * sizeof(double)));
  }
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int mom_product(Ctx &c, const double *x, double *y, bool jacobi);

// The momentum (or monolithic) operator with its preconditioner as a LinOp;
// the Jacobi-preconditioned product is one fused kernel.
struct MomOp : LinOp {
  Ctx &c;
  int pc;
  MomOp(Ctx &cc, int precond) : c(cc), pc(precond) { n = cc.mom.n; }
  int apply(const double *x, double *y) override { return mom_matvec(c, x, y); }
  int apply_pa(const double *x, double *y, double *tmp) override {
    if (pc == COVOF_PRECOND_JACOBI) return mom_product(c, x, y, true);
    return LinOp::apply_pa(x, y, tmp);
  }
  int precond(const double *r, double *z) override { return mom_precond(c, pc, r, z); }
};

SeqSet seq_set(Ctx &c);

// assemble_preconditioner_blocks into M.diag and (full) the SSOR / Schur
// coefficients of M.bk
int mom_blocks(Ctx &c, bool full) {
  Mom &M = c.mom;
  const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
  if (full || M.cfg.mono) {
    const int st = mom_alloc_blocks(c);
    if (st) return st;
  }
  BlkOut o{};
  o.diag = M.diag;
  if (full) {
    o.uW = M.bk[0], o.uE = M.bk[1], o.uS = M.bk[2], o.uN = M.bk[3], o.uP = M.bk[4];
    o.vW = M.bk[5], o.vE = M.bk[6], o.vS = M.bk[7], o.vN = M.bk[8], o.vP = M.bk[9];
  }
  if (M.cfg.mono) {
    o.uaP2 = M.aP2[0], o.vaP2 = M.aP2[1];
    hipLaunchKernelGGL(k_fill, grid_of(nh), dim3(BLK), 0, c.stream, nh, 1.0, M.aP2[0]);
    hipLaunchKernelGGL(k_fill, grid_of(nh), dim3(BLK), 0, c.stream, nh, 1.0, M.aP2[1]);
  }
  hipLaunchKernelGGL(k_mom_blocks, grid_of(M.nuv), dim3(BLK), 0, c.stream, c.m, M.cfg, M.f, M.dt, M.nu, M.nuv,
                     M.u_nx, M.v_ny, o);
  if (M.cfg.mono) {
    hipLaunchKernelGGL(k_exchange2, dim3(1), dim3(1024), 0, c.stream, c.m, M.aP2[0], M.aP2[1]);
    double *s[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    if (full)
      for (int q = 0; q < 5; ++q) s[q] = M.bk[10 + q];
    hipLaunchKernelGGL(k_schur, grid_of((long)c.m.nx * c.m.ny), dim3(BLK), 0, c.stream, c.m,
                       (const double *)M.aP2[0], (const double *)M.aP2[1], s[0], s[1], s[2], s[3], s[4],
                       M.diag + M.nuv);
  }
  if (full) {
    // the SSOR coefficients in the sweeps' skewed order
    const SeqSet a = seq_set(c);
    for (int b = 0; b < (M.cfg.mono ? 3 : 2); ++b) {
      const SeqBlk &q = a.blk[b];
      // records a_P, a_W, a_E, a_S, a_N from the blocks' W, E, S, N, P arrays
      const int field[5] = {SR_W, SR_E, SR_S, SR_N, SR_P};
      for (int k = 0; k < 5; ++k)
        hipLaunchKernelGGL(k_to_skew, grid_of((long)q.ni * q.nj), dim3(BLK), 0, c.stream, q.ni, q.nj, q.S,
                           (const double *)M.bk[5 * b + k], q.rec, field[k]);
    }
  }
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

}  // namespace

int mom_reserve(Ctx &c) {
  const int st = mom_alloc(c);
  return st ? st : mom_alloc_blocks(c);
}

void mom_free(Ctx &c) {
  Mom &M = c.mom;
  for (double *p : {M.tu, M.tv, M.tp, M.diag, M.rhs, M.x}) hipFree(p);
  for (double *p : M.w) hipFree(p);
  for (double *p : M.bk) hipFree(p);
  for (double *p : M.aP2) hipFree(p);
  for (double *p : M.sx) hipFree(p);
  for (auto &blk : M.sk)
    for (double *p : blk) hipFree(p);
  for (auto &blk : M.hal)
    for (double *p : blk) hipFree(p);
  M = Mom();
}

int mom_bind(Ctx &c, const MomCfg &cfg, const MomFields &f, double dt) {
  const int *bc = c.m.bc;
  if (!(dt > 0.0) || !f.u || !f.v || !f.uo || !f.vo || !f.ru || !f.rv || !f.mu || !f.mn || !f.fx || !f.fy)
    return COVOF_BAD_ARGUMENT;
  if ((cfg.gear || cfg.ab_nm1) && (!f.un || !f.vn)) return COVOF_BAD_ARGUMENT;
  if (cfg.lin < 1 || cfg.lin > 3 || cfg.scheme < 1 || cfg.scheme > 3) return COVOF_BAD_ARGUMENT;
  // periodic sides come in pairs; the reference's v-row at j = ny+1 reads past
  // its array with a NEUMANN top, so that side is not supported
  if ((bc[0] == BC_PERIODIC) != (bc[1] == BC_PERIODIC) || (bc[2] == BC_PERIODIC) != (bc[3] == BC_PERIODIC) ||
      bc[3] == BC_NEUMANN)
    return COVOF_BAD_ARGUMENT;
  if (c.obs.enabled) return COVOF_BAD_ARGUMENT;  // Brinkman / solid-obstacle rows are not built
  const int st = mom_alloc(c);
  if (st) return st;
  Mom &M = c.mom;
  M.cfg = cfg;
  M.f = f;
  M.dt = dt;
  M.u_nx = c.m.nx + (c.m.perx ? 0 : 1);
  M.v_ny = c.m.ny + (c.m.pery ? 0 : 1);
  M.nu = (long)M.u_nx * c.m.ny;
  M.nuv = M.nu + (long)c.m.nx * M.v_ny;
  M.n = M.nuv + (cfg.mono ? (long)c.m.nx * c.m.ny : 0);
  M.bound = 1;
  return COVOF_OK;
}

int mom_diag(Ctx &c, double *diag) {
  Mom &M = c.mom;
  if (!M.bound) return COVOF_BAD_ARGUMENT;
  int st = mom_blocks(c, false);
  if (st) return st;
  if (diag != M.diag)
    COVOF_CHECK(hipMemcpyAsync(diag, M.diag, M.n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  return COVOF_OK;
}

int mom_rhs(Ctx &c, double *rhs) {
  Mom &M = c.mom;
  if (!M.bound) return COVOF_BAD_ARGUMENT;
  hipLaunchKernelGGL(k_mom_rhs, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.cfg, M.f, M.dt, M.nu, M.nuv, M.n,
                     M.u_nx, M.v_ny, rhs);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int mom_pack(Ctx &c, const double *u, const double *v, const double *p, double *x) {
  Mom &M = c.mom;
  if (!M.bound || (M.cfg.mono && !p)) return COVOF_BAD_ARGUMENT;
  hipLaunchKernelGGL(k_pack, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.nu, M.nuv, M.n, M.u_nx, u, v, p, x);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

namespace {
// unpack_vector(homogeneous) into the work fields, then the operator (with
// the Jacobi division when jacobi: M.diag must hold the diagonal).
int mom_product(Ctx &c, const double *x, double *y, bool jacobi) {
  Mom &M = c.mom;
  double *tp = M.cfg.mono ? M.tp : nullptr;
  hipLaunchKernelGGL(k_unpack, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.nu, M.nuv, M.n, M.u_nx, x, M.tu,
                     M.tv, tp);
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, M.cfg, M.tu, M.tv, 1, 1, tp);
  launch_tagged(c, COVOF_KERNEL_MOM_MATVEC, jacobi ? k_mom_matvec<true> : k_mom_matvec<false>, grid_of(M.n),
                dim3(BLK), c.m, M.cfg, M.f, M.dt, M.nu, M.nuv, M.n, M.u_nx, M.v_ny, (const double *)M.tu,
                (const double *)M.tv, (const double *)M.tp, (const double *)M.diag, y);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// the blocks of the sequential preconditioners, as the kernels take them
SeqSet seq_set(Ctx &c) {
  Mom &M = c.mom;
  SeqSet a{};
  const int ni[3] = {M.u_nx, c.m.nx, c.m.nx}, nj[3] = {c.m.ny, M.v_ny, c.m.ny};
  for (int b = 0; b < 3; ++b) {
    SeqBlk &q = a.blk[b];
    q.ni = ni[b], q.nj = nj[b], q.K = (nj[b] + 63) / 64, q.S = ni[b] + 63;
    q.rec = M.sk[b][0], q.xa = M.sk[b][1], q.xb = M.sk[b][2];
    q.hW = M.hal[b][0], q.hE = M.hal[b][1], q.hS = M.hal[b][2], q.hN = M.hal[b][3];
  }
  return a;
}

// one wave per strip of the largest of the launched blocks, at most SEQ_MAXW
inline dim3 seq_threads(const SeqSet &a, int first, int nb) {
  int w = 1;
  for (int b = first; b < first + nb; ++b) w = seq_waves(a.blk[b].nj) > w ? seq_waves(a.blk[b].nj) : w;
  return dim3(64 * w);
}
}  // namespace

int mom_matvec(Ctx &c, const double *x, double *y) {
  if (!c.mom.bound) return COVOF_BAD_ARGUMENT;
  return mom_product(c, x, y, false);
}

// apply_preconditioner (src/module_linear_solvers.f90:48-316) with the blocks
// last assembled: NONE copies, JACOBI divides by the diagonal, SSOR runs two
// symmetric sweeps per block, BLOCK_TRIANGULAR (monolithic) the pressure
// Schur solve then the velocity solves against its gradient (SSOR otherwise).
int mom_precond(Ctx &c, int precond, const double *r, double *z) {
  Mom &M = c.mom;
  if (precond == COVOF_PRECOND_JACOBI) {
    hipLaunchKernelGGL(k_mom_jacobi, grid_of(M.n), dim3(BLK), 0, c.stream, M.n, (const double *)M.diag, r, z);
  } else if (precond == COVOF_PRECOND_SSOR || precond == COVOF_PRECOND_BLOCK_TRIANGULAR) {
    const SeqSet a = seq_set(c);
    const long off[3] = {0, M.nu, M.nuv};
    auto cells = [&](int b) { return grid_of((long)a.blk[b].ni * a.blk[b].nj); };
    if (precond == COVOF_PRECOND_BLOCK_TRIANGULAR && M.cfg.mono) {
      // pressure: 12 sweeps at omega 1.25, mean removed, exchanged
      double *xp = M.sx[2];
      const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
      hipLaunchKernelGGL(k_seq_rhs, cells(2), dim3(BLK), 0, c.stream, c.m, a.blk[2], 0, r + off[2],
                         (const double *)nullptr);
      hipLaunchKernelGGL(k_seq_ssor, dim3(1), seq_threads(a, 2, 1), 0, c.stream, c.m, a, 2, 12, 1.25);
      hipLaunchKernelGGL(k_fill, grid_of(nh), dim3(BLK), 0, c.stream, nh, 0.0, xp);
      hipLaunchKernelGGL(k_from_skew, cells(2), dim3(BLK), 0, c.stream, c.m, a.blk[2], (const double *)a.blk[2].xa,
                         xp, 1);
      hipLaunchKernelGGL(k_bt_mean, dim3(1), dim3(1024), 0, c.stream, c.m, xp, z + off[2]);
      // velocity: 4 sweeps each against rhs - grad p
      for (int b = 0; b < 2; ++b)
        hipLaunchKernelGGL(k_seq_rhs, cells(b), dim3(BLK), 0, c.stream, c.m, a.blk[b], b + 1, r + off[b],
                           (const double *)xp);
      hipLaunchKernelGGL(k_seq_ssor, dim3(2), seq_threads(a, 0, 2), 0, c.stream, c.m, a, 0, 4, 1.0);
      for (int b = 0; b < 2; ++b)
        hipLaunchKernelGGL(k_from_skew, cells(b), dim3(BLK), 0, c.stream, c.m, a.blk[b],
                           (const double *)a.blk[b].xa, z + off[b], 0);
    } else {
      const int nb = M.cfg.mono ? 3 : 2;
      for (int b = 0; b < nb; ++b)
        hipLaunchKernelGGL(k_seq_rhs, cells(b), dim3(BLK), 0, c.stream, c.m, a.blk[b], 0, r + off[b],
                           (const double *)nullptr);
      hipLaunchKernelGGL(k_seq_ssor, dim3(nb), seq_threads(a, 0, nb), 0, c.stream, c.m, a, 0, 2, 1.0);
      for (int b = 0; b < nb; ++b)
        hipLaunchKernelGGL(k_from_skew, cells(b), dim3(BLK), 0, c.stream, c.m, a.blk[b],
                           (const double *)a.blk[b].xa, z + off[b], 0);
    }
  } else {
    if (z != r) COVOF_CHECK(hipMemcpyAsync(z, r, M.n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return COVOF_OK;
  }
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

static bool precond_ok(int precond) {
  return precond == COVOF_PRECOND_NONE || precond == COVOF_PRECOND_JACOBI || precond == COVOF_PRECOND_SSOR ||
         precond == COVOF_PRECOND_BLOCK_TRIANGULAR;
}

// assemble_preconditioner_blocks of the bound fields, then apply_preconditioner
int mom_apply_precond(Ctx &c, int precond, const double *r, double *z) {
  if (!c.mom.bound || !precond_ok(precond)) return COVOF_BAD_ARGUMENT;
  const int st = mom_blocks(c, precond == COVOF_PRECOND_SSOR || precond == COVOF_PRECOND_BLOCK_TRIANGULAR);
  return st ? st : mom_precond(c, precond, r, z);
}

// solve_linear_system on the momentum (or monolithic) operator, the
// preconditioner from the blocks of the bound fields.  b, x: n unknowns, not
// a work vector.
int mom_solve(Ctx &c, const double *b, double *x, int lin_solver, int precond, double tol, int maxit, int *iters,
              double *relres) {
  Mom &M = c.mom;
  if (!M.bound) return COVOF_BAD_ARGUMENT;
  if ((lin_solver != COVOF_LINEAR_BICGSTAB2 && lin_solver != COVOF_LINEAR_CG) || !precond_ok(precond) ||
      maxit < 0)
    return COVOF_BAD_ARGUMENT;
  int st = mom_blocks(c, precond == COVOF_PRECOND_SSOR || precond == COVOF_PRECOND_BLOCK_TRIANGULAR);
  if (st) return st;
  MomOp op(c, precond);
  if (lin_solver == COVOF_LINEAR_CG) return krylov_cg(c, op, b, x, M.w, tol, maxit, iters, relres);
  return krylov_bicgstab2(c, op, b, x, M.w, tol, maxit, iters, relres);
}

// sync_u, sync_v and apply_boundary_conditions of one rank
// (src/module_mpi_comm.f90:733-775, :290-537) on halo fields u, v.
int velocity_bc(Ctx &c, const MomCfg &cfg, double *u, double *v, int homogeneous) {
  const int *bc = c.m.bc;
  if ((bc[0] == BC_PERIODIC) != (bc[1] == BC_PERIODIC) || (bc[2] == BC_PERIODIC) != (bc[3] == BC_PERIODIC))
    return COVOF_BAD_ARGUMENT;
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, cfg, u, v, homogeneous ? 1 : 0, 1,
                     (double *)nullptr);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// The momentum block of sim_solve_standard_projection (:2454-2469): the
// current u, v are the initial guess and the operator's fields_old; the
// solution is unpacked into them with the inhomogeneous BC, then synced and
// the BC applied once more, as the reference does.
int mom_predictor(Ctx &c, double *u, double *v, int lin_solver, int precond, double tol, int maxit, int *iters,
                  double *relres) {
  Mom &M = c.mom;
  if (!M.bound || M.cfg.mono) return COVOF_BAD_ARGUMENT;
  int st = mom_rhs(c, M.rhs);
  if (!st) st = mom_pack(c, u, v, nullptr, M.x);
  if (st) return st;
  const int sst = mom_solve(c, M.rhs, M.x, lin_solver, precond, tol, maxit, iters, relres);
  if (sst == COVOF_BAD_ARGUMENT || sst == COVOF_HIP_ERROR) return sst;
  hipLaunchKernelGGL(k_unpack, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.nu, M.nuv, M.n, M.u_nx,
                     (const double *)M.x, u, v, (double *)nullptr);
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, M.cfg, u, v, 0, 1, (double *)nullptr);
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, M.cfg, u, v, 0, 1, (double *)nullptr);
  COVOF_CHECK_LAUNCH();
  return sst;
}

// The monolithic branch of the time step (src/module_simulation.f90:553-583,
// one Uzawa pass): blocks, right-hand side with zero continuity rows, pack
// (u, v, p: the initial guess), solve, unpack_vector(.false.) into u, v, p
// (pressure exchanged, velocity exchanged and its BC applied).
int mono_step(Ctx &c, double *u, double *v, double *p, int lin_solver, int precond, double tol, int maxit,
              int *iters, double *relres) {
  Mom &M = c.mom;
  if (!M.bound || !M.cfg.mono || !p) return COVOF_BAD_ARGUMENT;
  int st = mom_rhs(c, M.rhs);
  if (!st) st = mom_pack(c, u, v, p, M.x);
  if (st) return st;
  const int sst = mom_solve(c, M.rhs, M.x, lin_solver, precond, tol, maxit, iters, relres);
  if (sst == COVOF_BAD_ARGUMENT || sst == COVOF_HIP_ERROR) return sst;
  hipLaunchKernelGGL(k_unpack, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.nu, M.nuv, M.n, M.u_nx,
                     (const double *)M.x, u, v, p);
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, M.cfg, u, v, 0, 1, p);
  COVOF_CHECK_LAUNCH();
  return sst;
}

}  // namespace covof
